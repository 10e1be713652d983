"""The memory network (nrc_ensemble_set_memnet, nrc_z2_memnet_eval, nrc_z2_memnet_grad): a
pointwise network over each site's two decayed memories of its own past field, an exact
importance sampler for the Z2 field of src/lattice/lattice.C:81-86.  The fields the GPU draws,
replayed through the oracle (oracle/memnet.py), give the same determinants and the oracle's
log ratios; the evaluation kernel gives the oracle's ratios on the plain sampler's fields;
the gradient kernel gives the oracle's gradient; training on a plain sample raises the
effective-sample objective on held-out configurations, and the reweighted correlator agrees
with the plain sampler's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import nrcps_amd as na

    c = na.Context(0)
    yield c
    c.close()


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(35).astype(np.float32) * np.float32(scale)
    p[34] = np.float32(40.0 * scale)  # a tilt of logit ~0.04 beside the memory terms
    return p


def _plain_phis(orc, seed, config, T, V):
    return np.stack([orc.bits_to_phi(orc.z2_bits(seed, config, t, V), V) for t in range(T)])


@pytest.mark.parametrize("prec,conj", [("f64", False), ("f32", True)])
def test_memnet_fields_replayed_by_oracle(ctx, orc, prec, conj):
    import nrcps_amd as na
    from oracle import memnet

    L, N, T, C, seed, c0 = 8, 7, 6, 4, 93, 500
    V = L ** 3
    print(f"[gpu] memory network {prec} conj={conj} L={L} N={N} T={T}", flush=True)
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    e = na.Ensemble(ctx, L, moms, kin, 0.4, C,
                    prec=na.PREC_F64 if prec == "f64" else na.PREC_F32)
    e.set_symmetry(conj)
    p = _params(5, 8.0)
    e.set_memnet(p)
    la, ph, _ = e.run(seed, c0, T)
    lr = e.log_ratio(T)
    e.close()
    tol = 1e-10 if prec == "f64" else 5e-6
    for c in range(C):
        phi, want = memnet.sample(p, seed, c0 + c, T, V)
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, field="array", phi=phi)
        assert np.abs(la[:, c] - ref["logabs"]).max() < tol
        assert np.abs(ph[:, c] - ref["phase"]).max() < tol
        assert np.allclose(lr[:, c], want, rtol=1e-11, atol=1e-9), c
    assert np.all(lr[0] == 0) and np.all(lr[1:] != 0)


@pytest.mark.parametrize("L,T,c0", [(8, 6, 40), (24, 3, 1 << 22), (32, 2, 7)])
def test_memnet_eval_matches_oracle(ctx, orc, L, T, c0):
    """The evaluation kernel on the plain sampler's fields (one or two words per thread, 512
    threads at L = 32)."""
    from nrcps_amd import proposal as pr
    from oracle import memnet

    V, seed, C = L ** 3, 14, 3
    p = _params(7, 4.0)
    print(f"[gpu] memory network eval L={L} T={T} configs {c0}..", flush=True)
    got = pr.memnet_log_ratio(ctx, seed, c0, C, T, V, p)
    for c in range(C):
        want = memnet.log_ratio(p, _plain_phis(orc, seed, c0 + c, T, V))
        assert np.allclose(got[:, c], want, rtol=1e-11, atol=1e-8), c


@pytest.mark.parametrize("L,T", [(8, 6), (24, 3)])
def test_memnet_gradient_matches_oracle(ctx, orc, L, T):
    from nrcps_amd import proposal as pr
    from oracle import memnet

    V, seed, C, c0 = L ** 3, 15, 3, 60
    p = _params(9, 2.0)
    rng = np.random.default_rng(L)
    G = rng.random((T, C))
    G[1, 0] = 0.0  # a skipped slice
    print(f"[gpu] memory network gradient L={L} T={T}", flush=True)
    got = pr.memnet_gradient(ctx, seed, c0, C, T, V, p, G)
    want = sum(memnet.grad_log_q(p, _plain_phis(orc, seed, c0 + c, T, V), G[:, c])
               for c in range(C))
    assert np.abs(got - want).max() <= 1e-4 * np.abs(want).max()


def test_memnet_trained_raises_the_objective_and_is_unbiased(ctx, orc):
    """Trained on a plain sample at L = 8, the network's effective-sample objective (the mean
    log effective-sample fraction) on the held-out half exceeds the plain sampler's, and a
    fresh run with it gives the plain sampler's correlator within errors."""
    import nrcps_amd as na
    from nrcps_amd import proposal as pr

    L, N, T, seed = 8, 7, 16, 31
    V = L ** 3
    ts = [4, 8, 12]
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    e = na.Ensemble(ctx, L, moms, kin, 0.4, 2048, prec=na.PREC_F32)
    la, _, _ = e.run(seed, 0, T)
    p, rec = pr.fit_memnet(ctx, la[:, :, 0], seed, 0, T, V, ts, pr.memnet_init(0.0, 3),
                           iters=40, step=0.05)
    print(f"[gpu] memory network trained: validation objective {rec['validation_objective']:.4f}"
          f" at iteration {rec['best_iteration']}, plain {rec['validation_objective_plain']:.4f}",
          flush=True)
    assert rec["validation_objective"] > rec["validation_objective_plain"]
    # the reweighted correlator of a fresh run against the plain one (other configurations)
    C2 = 4096
    f = na.Ensemble(ctx, L, moms, kin, 0.4, C2, prec=na.PREC_F32)
    f.set_memnet(p)
    lq, phq, _ = f.run(seed, 10000, T)
    lr = f.log_ratio(T)
    f.close()
    lp, php, _ = e.run(seed, 20000, T)
    e.close()
    for t in ts:
        wq = np.exp(2 * lq[t, :, 0] + lr[t]) * (phq[t, :, 0] ** 2).real
        wp = np.exp(2 * lp[t, :, 0]) * (php[t, :, 0] ** 2).real
        mq, sq = wq.mean(), wq.std() / np.sqrt(wq.size)
        mp, sp = wp.mean(), wp.std() / np.sqrt(wp.size)
        assert abs(mq - mp) < 4 * np.hypot(sq, sp), (t, mq, mp, sq, sp)
