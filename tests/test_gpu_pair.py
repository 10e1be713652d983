"""The pair-sink matrix nrc_pair_overlap (TwoBody::Run) through the C ABI against the FP64
oracle (oracle pair_overlap, pinned by brute force): FP64 and FP32 inputs with FP64 sums,
and split-BF16x3 products on MFMA."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import nrcps_amd as na

    c = na.Context(0)
    yield c
    c.close()


def _rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _case(orc, L, n1, n2, ncfg, seed, bc):
    rng = np.random.default_rng(seed)
    V = L ** 3
    a = _rand_c(rng, (ncfg, n1, V))
    b = _rand_c(rng, (ncfg, n2, V))
    w = orc.two_body_wavefunc(L, orc.WAVEFUNC_PAIR1, 0.3, cutoff_sq=3 * np.pi ** 2, bc=bc)
    opp = orc.opposite_parity_index(L, bc)
    return a, b, w, opp


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("L,n1,n2,ncfg,bc", [(8, 5, 7, 3, (0, 0, 0)), (6, 3, 3, 2, (1, 0, 1)),
                                             (4, 1, 1, 1, (0, 0, 0))])
def test_pair_overlap_dense(ctx, orc, mode, L, n1, n2, ncfg, bc):
    print(f"[gpu] pair overlap mode={mode} L={L} n1={n1} n2={n2}", flush=True)
    a, b, w, opp = _case(orc, L, n1, n2, ncfg, L * 10 + n1, bc)
    dt = np.complex128 if mode == 0 else np.complex64
    a, b = a.astype(dt), b.astype(dt)
    M = ctx.array((ncfg, n1, n2), np.complex128)
    ctx.pair_overlap(ctx.to_device(a), ctx.to_device(b), ctx.to_device(w),
                     ctx.to_device(opp.astype(np.int32)), M, n1, n2, L ** 3, ncfg, mode)
    got = M.download()
    for c in range(ncfg):
        ref = orc.pair_overlap(a[c].astype(np.complex128), b[c].astype(np.complex128), w, opp)
        np.testing.assert_allclose(got[c], ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("L,n1,n2,ncfg,bc", [(32, 35, 35, 2, (0, 0, 0)), (16, 7, 7, 3, (0, 0, 0)),
                                             (8, 64, 64, 1, (0, 0, 0)), (24, 14, 19, 2, (0, 0, 0)),
                                             (16, 9, 5, 2, (1, 1, 1)), (8, 3, 17, 2, (0, 1, 1))])
def test_pair_overlap_bf16x3_mfma(ctx, orc, L, n1, n2, ncfg, bc):
    """Each product carries ~2^-17 relative error; unit vectors times values with 16
    significant bits are exact.  Periodic boundaries put the mirrored runs at odd sites,
    antiperiodic ones at 16-byte aligned sites (both load paths)."""
    print(f"[gpu] pair overlap bf16x3 L={L} n1={n1} n2={n2} bc={bc}", flush=True)
    a, b, w, opp = _case(orc, L, n1, n2, ncfg, L + n1, bc)
    a, b = a.astype(np.complex64), b.astype(np.complex64)
    wf = w.astype(np.float32).astype(np.float64)
    M = ctx.array((ncfg, n1, n2), np.complex128)
    ctx.pair_overlap(ctx.to_device(a), ctx.to_device(b), ctx.to_device(wf),
                     ctx.to_device(opp.astype(np.int32)), M, n1, n2, L ** 3, ncfg, 2)
    got = M.download()
    for c in range(ncfg):
        ref = orc.pair_overlap(a[c].astype(np.complex128), b[c].astype(np.complex128), wf, opp)
        scale = np.sqrt(np.sum(wf ** 2)) * 2.0  # random-sign sum of V weighted products
        assert np.abs(got[c] - ref).max() < 2e-5 * scale, np.abs(got[c] - ref).max()
    # unit rows at chosen sites pick w(k) psi2_t(opp(k)) exactly
    V = L ** 3
    rng = np.random.default_rng(5)
    sites = rng.choice(V, n1, replace=False)
    e = np.zeros((ncfg, n1, V), np.complex64)
    e[:, np.arange(n1), sites] = 1
    q = (np.round(_rand_c(rng, (ncfg, n2, V)) * 256) / 256).astype(np.complex64)
    ones = np.ones(V)
    ctx.pair_overlap(ctx.to_device(e), ctx.to_device(q), ctx.to_device(ones),
                     ctx.to_device(opp.astype(np.int32)), M, n1, n2, V, ncfg, 2)
    got = M.download()
    assert np.array_equal(got, q[:, :, opp[sites]].transpose(0, 2, 1).astype(np.complex128))


@pytest.mark.parametrize("prec", [1, 0])
def test_ensemble_pair_matches_oracle(ctx, orc, prec):
    """The pair-sink measurement of the batched slice loop (SlaterDet3's 2N determinants per
    slice) against the oracle's run on the same Philox fields: FP64 to 1e-9, FP32 (split-
    BF16x3 pair products) to 1e-4 in log|det|."""
    import nrcps_amd as na

    L, N, T, C, seed = 8, 7, 6, 3, 23
    print(f"[gpu] ensemble pair sinks prec={prec} L={L} N={N} T={T}", flush=True)
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    w = orc.two_body_wavefunc(L, orc.WAVEFUNC_TGND, 0.176776695, cutoff_sq=np.pi ** 2)
    opp = orc.opposite_parity_index(L)
    ens = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=prec, split=na.SPLIT_KSPACE,
                      pair_w=w, pair_opp=opp)
    la, ph, _ = ens.run(seed, 0, T)
    assert la.shape == (T, C, 2 * N)
    tol = 1e-9 if prec == 1 else 1e-4
    for c in range(C):
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, seed=seed, config=c, pair_w=w,
                             pair_opp=opp)
        both_inf = np.isneginf(la[:, c, :]) & np.isneginf(ref["logabs"])  # exact zeros, t = 0
        assert np.all(np.isneginf(la[:, c, :]) == np.isneginf(ref["logabs"]))
        err = np.abs(la[:, c, :][~both_inf] - ref["logabs"][~both_inf]).max()
        perr = np.abs(ph[:, c, :] - ref["phase"]).max()
        print(f"[gpu] config {c}: max |d log|det|| {err:.2e}, phase {perr:.2e}", flush=True)
        assert err < tol and perr < 50 * tol
    ens.close()


@pytest.mark.parametrize("prec,N,L,signed", [(1, 7, 8, False), (1, 9, 8, False),
                                             (0, 7, 8, False), (0, 35, 16, False),
                                             (0, 9, 8, True), (0, 35, 16, True)])
def test_ensemble_pair_conj_partners_match_oracle(ctx, orc, prec, N, L, signed):
    """Pair sinks with conjugate partners: the pair matrices come from two products over the
    stored volumes (mirrored sites, and the same sites conjugated; in FP32 one pass computes
    both); N = 9 leaves two orbitals of a partial shell without partner, N = 35 (19 volumes)
    is the pair-sink benchmark's orbital set.  `signed` multiplies the wave function by
    cos(2 pi n_x / L), even under p -> -p and negative on half the lattice, so the FP32 pass
    carries the sign of w on its products.  Against the oracle's run, which evolves every
    orbital: FP64 1e-9, FP32 1e-4 in log|det| (5e-4 for the 35 x 35 minors, whose pivots
    amplify the ~2^-17 relative error of the split products; the every-orbital FP32 run is
    checked to the same bound beside it)."""
    import nrcps_amd as na

    T, C, seed = 6, 3, 29
    print(f"[gpu] ensemble pair sinks, conjugate partners, prec={prec} N={N} L={L} "
          f"signed={signed}", flush=True)
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    w = orc.two_body_wavefunc(L, orc.WAVEFUNC_TGND, 0.176776695, cutoff_sq=np.pi ** 2)
    if signed:
        nx = np.arange(L ** 3) // (L * L)
        w = w * np.cos(2 * np.pi * nx / L)
        assert (w < 0).any()
    opp = orc.opposite_parity_index(L)
    ens = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=prec, split=na.SPLIT_KSPACE,
                      pair_w=w, pair_opp=opp)
    ens.set_symmetry(True)
    assert ens.volumes_per_config() == {7: 4, 9: 6, 35: 19}[N]
    la, ph, _ = ens.run(seed, 0, T)
    tol = 1e-9 if prec == 1 else (1e-4 if N < 20 else 5e-4)
    full = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=prec, split=na.SPLIT_KSPACE,
                       pair_w=w, pair_opp=opp)
    la_full, _, _ = full.run(seed, 0, T)
    full.close()
    for c in range(C):
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, seed=seed, config=c, pair_w=w,
                             pair_opp=opp)
        fin = np.isfinite(ref["logabs"])
        efull = np.abs(la_full[:, c, :][fin] - ref["logabs"][fin]).max()
        print(f"[gpu] config {c}: every orbital evolved, max |d log|det|| {efull:.2e}",
              flush=True)
        assert efull < tol
        both_inf = np.isneginf(la[:, c, :]) & np.isneginf(ref["logabs"])
        assert np.all(np.isneginf(la[:, c, :]) == np.isneginf(ref["logabs"]))
        err = np.abs(la[:, c, :][~both_inf] - ref["logabs"][~both_inf]).max()
        perr = np.abs(ph[:, c, :] - ref["phase"]).max()
        print(f"[gpu] config {c}: max |d log|det|| {err:.2e}, phase {perr:.2e}", flush=True)
        assert err < tol and perr < 50 * tol
    ens.close()


def test_ensemble_pair_conj_partners_refused_for_twisted_mirror(ctx, orc):
    """Antiperiodic boundaries shift the opposite-parity site off -p: no partners."""
    import nrcps_amd as na

    L, N = 8, 7
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    w = orc.two_body_wavefunc(L, orc.WAVEFUNC_PAIR1, 0.3, cutoff_sq=3 * np.pi ** 2, bc=(1, 0, 0))
    opp = orc.opposite_parity_index(L, (1, 0, 0))
    ens = na.Ensemble(ctx, L, moms, kin, 0.4, 2, prec=1, split=na.SPLIT_KSPACE, pair_w=w,
                      pair_opp=opp)
    with pytest.raises(na.NrcError):
        ens.set_symmetry(True)
    ens.close()
