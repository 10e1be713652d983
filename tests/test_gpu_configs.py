"""The workload shapes beyond the headline, on the GPU against the oracle or the FP64 path:
  * config 1 (L=12, N=1, k=0, T=32, x-space split, FP64, 500 configurations): every
    configuration's S(t) against the oracle and the Monte Carlo <S(t)^2> against the exact
    transfer-matrix expectation (tests/free_gas/main.C:71-73 is the reference's x-space shape);
  * config 4 (batched uneven shapes): FP32 groups at L = 24 and 32 with conjugate partners,
    N in {7, 19, 27, 33}, T = 96 in one captured graph, against the same configurations on the
    FP64 path;
  * config 5 (L=24, N=14, the partial shell: 7 of the 12 |n|^2 = 2 momenta in the
    reference's tie order): the ensemble with conjugate partners against the oracle, and one
    run of the learned sampler whose sampled fields are replayed through the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import nrcps_amd as na

    c = na.Context(0)
    yield c
    c.close()


def _bits_to_phi(words, V):
    """Philox / proposal words [..., ceil(V/32)] -> phi [..., V] in {-1, +1} (int8)."""
    w = np.asarray(words, dtype=np.uint32)
    bits = (w[..., :, None] >> np.arange(32, dtype=np.uint32)) & 1
    return (2 * bits.reshape(w.shape[:-1] + (-1,))[..., :V] - 1).astype(np.int8)


def test_config1_xspace_fp64_vs_oracle_and_exact_expectation(ctx, orc):
    import nrcps_amd as na

    L, T, A, ncfg = 12, 32, 0.4, 500
    print(f"[gpu] config 1: L={L} N=1 T={T} x-space split FP64, {ncfg} configurations",
          flush=True)
    kin = orc.kinetic_exp(L, 0.05, 1.0)
    moms = np.zeros((1, 3), dtype=np.int32)
    e = na.Ensemble(ctx, L, moms, kin, A, ncfg, prec=na.PREC_F64, split=na.SPLIT_XSPACE)
    la, ph, _ = e.run(7012, 0, T)
    e.close()
    for c in range(0, ncfg, 25):  # every 25th configuration against the oracle
        ref = orc.run_config(L, 1, T, moms, kin, A=A, split=orc.SPLIT_XSPACE, seed=7012,
                             config=c)
        assert np.abs(la[:, c, 0] - ref["logabs"][:, 0]).max() < 1e-12
        assert np.abs(ph[:, c, 0] - ref["phase"][:, 0]).max() < 1e-12
    s2 = (np.exp(2 * la[:, :, 0]) * ph[:, :, 0] ** 2).real
    exact = orc.exact_c1(L, T, kin, np.ones(L ** 3), A)
    mean, err = s2.mean(1), s2.std(1, ddof=1) / np.sqrt(ncfg)
    print(f"[gpu]   max |<S^2> - exact| / sigma = {np.max(np.abs(mean - exact)[1:] / err[1:]):.2f}",
          flush=True)
    assert np.all(np.abs(mean[1:] - exact[1:]) < 4 * err[1:] + 1e-12)
    assert mean[0] == pytest.approx(1.0, abs=1e-12)


def test_config4_batch_l24_l32_groups_vs_fp64_path(ctx, orc):
    """FP32 groups (FP64 x-chain, conjugate partners) in one captured batch at T = 96, each
    against the FP64 path (every orbital) on the same configurations."""
    import nrcps_amd as na

    shapes = [(24, 7, 8), (32, 19, 4), (32, 33, 4), (24, 33, 4), (32, 7, 6), (24, 27, 4)]
    T, seed, c0 = 96, 4096, [100, 200, 300, 400, 500, 600]
    print(f"[gpu] config 4: batch of {len(shapes)} groups at L in {{24, 32}}, T={T}", flush=True)
    tables = {L: orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5) for L in (24, 32)}
    groups = []
    for L, N, C in shapes:
        g = na.Ensemble(ctx, L, orc.lowest_momenta(L, N), tables[L], 0.4, C, prec=na.PREC_F32)
        g.set_symmetry(True)
        groups.append(g)
    b = na.Batch(ctx, groups)
    b.run(seed, c0, T, graph=True)
    worst = 0.0
    for g, (L, N, C), c in zip(groups, shapes, c0):
        la, ph, _ = g.results(T)
        r = na.Ensemble(ctx, L, orc.lowest_momenta(L, N), tables[L], 0.4, C, prec=na.PREC_F64)
        la64, ph64, _ = r.run(seed, c, T)
        r.close()
        d = np.abs(la - la64).max()
        worst = max(worst, d)
        print(f"[gpu]   L={L} N={N}: max |log|det| FP32 - FP64| = {d:.2e}", flush=True)
        assert d < 5e-6
        assert np.abs(ph - ph64).max() < 1e-6
    b.close()
    for g in groups:
        g.close()


def test_config5_l24_n14_partial_shell_vs_oracle(ctx, orc):
    """The partial shell (N = 14) with conjugate partners (9 volumes: three unpaired momenta),
    FP64 against the oracle (which evolves every orbital) and FP32 against FP64."""
    import nrcps_amd as na

    L, N, T, C, seed = 24, 14, 10, 3, 7012
    print(f"[gpu] config 5: L={L} N={N} T={T}", flush=True)
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    e = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F64)
    e.set_symmetry(True)
    assert e.volumes_per_config() == 9
    la, ph, _ = e.run(seed, 50, T)
    e.close()
    for c in range(C):
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, seed=seed, config=50 + c)
        assert np.abs(la[:, c] - ref["logabs"]).max() < 1e-10
        assert np.abs(ph[:, c] - ref["phase"]).max() < 1e-10
    f = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F32)
    f.set_symmetry(True)
    la32, _, _ = f.run(seed, 50, T)
    f.close()
    assert np.abs(la32 - la).max() < 2e-6


def test_config5_learned_sampler_l24_fields_replayed_by_oracle(ctx, orc):
    """One run of the learned sampler at L = 24, N = 14 (FP32 ensemble, conjugate partners):
    the fields it draws slice by slice, replayed through the oracle, give the same
    determinants; the accumulated log ratio is finite and, for the first slice, the oracle's
    (oracle/convnet.py) exact ratio of the sampled field."""
    import nrcps_amd as na
    from nrcps_amd import proposal as pr
    from oracle import convnet as cn

    L, N, T, B, seed, config0 = 24, 14, 6, 3, 77, 9000
    print(f"[gpu] config 5: learned sampler L={L} N={N} T={T} B={B}", flush=True)
    V = L ** 3
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    ens = na.Ensemble(ctx, L, moms, kin, 0.4, B, prec=na.PREC_F32)
    ens.set_symmetry(True)
    params = pr.init_params(np.random.default_rng(3), out_scale=0.3)
    prop = pr.Proposal(ctx, L, B, params)
    mu, sd = np.zeros(T), np.ones(T)
    sampler = pr.LearnedSampler(ctx, prop, ens, mu, sd)
    # the sampler's loop, keeping each slice's words
    ens.begin(T)
    la_dev, _ = ens.results_dev()
    logr = sampler.logr
    ctx.lib.nrc_memset(ctx.h, logr.ptr, 0, logr.nbytes)
    words, prev, logr0, lg0 = [], None, None, None
    logits = ctx.array((B, V), np.float32)
    for t in range(T):
        ens.measure(t)
        pr.norm_logw(ctx, la_dev + t * B * ens.nm * 8, ens.nm, B, 0.0, 1.0, sampler.u)
        cur = sampler.words[t & 1]
        if t == 0:
            prop.forward(None, sampler.u, B, logits)
            lg0 = logits.download().copy()
        prop.sample(prev, sampler.u, B, seed, config0, t, cur, logr)
        ens.evolve(t, cur)
        words.append(cur.download().copy())
        if t == 0:
            logr0 = logr.download().copy()
        prev = cur
    la, ph, _ = ens.results(T)
    lr = logr.download()
    prop.close()
    ens.close()
    phi = _bits_to_phi(np.stack(words), V)  # [T][B][V]
    for c in range(B):
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, field="array", phi=phi[:, c])
        # the stepping kernels are plain FP32: FP32 evolution error over T = 6 slices
        assert np.abs(la[:, c] - ref["logabs"]).max() < 5e-6
        assert np.abs(ph[:, c] - ref["phase"]).max() < 5e-6
    assert np.all(np.isfinite(lr))
    # first slice: the exact log ratio of the drawn field under the proposal's logits
    for c in range(B):
        r, _ = cn.sample_ratio(lg0[c], phi[0, c] > 0)
        assert logr0[c] == pytest.approx(r.sum(), rel=1e-12, abs=1e-12)
    print(f"[gpu]   sampled fraction of +1: {(phi == 1).mean():.4f}; log r after T: "
          f"{lr.mean():.3f}", flush=True)


def test_config2_full_size_fp64(ctx, orc):
    """Config 2 at its full size (L=16, N=7, T=64, 2048 configurations, FP64): the Philox
    fields of every configuration bit-exact against the oracle at three slices, every 64th
    configuration's log|det| and phase against the oracle to 1e-12, and the correlator's
    sample mean over all 2048 consistent with the sampled configurations' own (a check that
    the batch carries no configuration-dependent offset)."""
    import nrcps_amd as na

    L, N, T, C, seed = 16, 7, 64, 2048, 7012
    V = L ** 3
    print(f"[gpu] config 2: L={L} N={N} T={T} C={C} FP64", flush=True)
    words = ctx.array((C, V // 32), np.uint32)
    for t in (0, 31, 63):
        ctx.lib.nrc_z2_bits(ctx.h, seed, 0, C, t, 0, V, words.ptr)
        got = words.download()
        ref = np.stack([orc.z2_bits(seed, c, t, V) for c in range(C)])
        assert np.array_equal(got, ref)
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    e = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F64)
    la, ph, _ = e.run(seed, 0, T)
    e.close()
    sample = range(0, C, 64)
    for c in sample:
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, seed=seed, config=c)
        assert np.abs(la[:, c] - ref["logabs"]).max() < 1e-12
        assert np.abs(ph[:, c] - ref["phase"]).max() < 1e-12
    d2 = (np.exp(2 * la[:, :, 0]) * ph[:, :, 0] ** 2).real
    full, sub = d2.mean(axis=1), d2[:, list(sample)].mean(axis=1)
    se = d2[:, list(sample)].std(axis=1, ddof=1) / np.sqrt(len(sample))
    assert np.all(np.isfinite(full)) and full[0] == pytest.approx(1.0, abs=1e-12)
    assert np.all(np.abs(full[1:] - sub[1:]) < 5 * se[1:] + 1e-300)
