"""The precision form of the workload (L=32, N=33 plane waves, T=128, tau=0.05, A=0.4, 1024
configurations): FP32 evolution must hold the correlator C(t) = <det(t)^2> within 1e-6
relative of the FP64 path on the same Philox fields at every t.  The FP64 reference evolves
every orbital with the FP64 kernels (the oracle's algorithm; those kernels are held to the
oracle at 1e-12 elsewhere).  Order of measurement and evolution as in
production/slater_det/main.C:201-220."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L, N, T, NCONF, SEED = 32, 33, 128, 1024, 7012
TOL_C = 1e-6  # relative error of C(t), every t


@pytest.fixture(scope="module")
def ctx():
    import nrcps_amd as na

    c = na.Context(0)
    yield c
    c.close()


def _lowest(L, N):
    f = np.where(np.arange(L) > L // 2, np.arange(L) - L, np.arange(L))
    pts = [(a, b, c) for a in range(L) for b in range(L) for c in range(L)]
    q = [f[a] ** 2 + f[b] ** 2 + f[c] ** 2 for a, b, c in pts]
    order = sorted(range(len(pts)), key=lambda i: q[i])  # ties keep enumeration order
    return np.array([pts[i] for i in order[:N]], dtype=np.int32)


def _xi_half(L, tau=0.05, M=1.0):
    n = np.arange(L)
    s = np.sin(np.pi * n / L) ** 2
    eps = 2.0 * (s[:, None, None] + s[None, :, None] + s[None, None, :]) / M
    return (1.0 + 0.5 * tau * eps).reshape(-1)


def _corr(la, ph):
    """log-scale and scaled C(t) = <det^2> over configurations (full-matrix results)."""
    la2 = 2.0 * la[:, :, 0]
    m = la2.max(axis=1, keepdims=True)
    return m[:, 0], (np.exp(la2 - m) * ph[:, :, 0] ** 2).mean(axis=1)


@pytest.fixture(scope="module")
def fp64_reference(ctx):
    import nrcps_amd as na

    print(f"[gpu] FP64 reference: L={L} N={N} T={T} {NCONF} configurations", flush=True)
    e = na.Ensemble(ctx, L, _lowest(L, N), _xi_half(L), 0.4, NCONF, prec=na.PREC_F64)
    la, ph, _ = e.run(SEED, 0, T)
    e.close()
    return la, ph


def _fp32_run(ctx, conj, acc):
    import nrcps_amd as na

    e = na.Ensemble(ctx, L, _lowest(L, N), _xi_half(L), 0.4, NCONF, prec=na.PREC_F32)
    e.set_symmetry(conj)
    e.set_accuracy(acc)
    la, ph, _ = e.run(SEED, 0, T)
    e.close()
    return la, ph


@pytest.mark.parametrize("conj", [True, False])
def test_config3_correlator_within_1e6_of_fp64(ctx, fp64_reference, conj):
    """The FP32 path with the FP64 x-chain (the default accuracy) holds every C(t)."""
    import nrcps_amd as na

    print(f"[gpu] FP32 (x-chain in FP64) conj={conj}", flush=True)
    la, ph = _fp32_run(ctx, conj, na.ACC_X64)
    m64, c64 = _corr(*fp64_reference)
    m32, c32 = _corr(la, ph)
    rel = np.abs(c32 * np.exp(m32 - m64) - c64) / np.abs(c64)
    print(f"[gpu]   max rel err of C(t) {rel.max():.2e} at t={rel.argmax()}, median "
          f"{np.median(rel):.2e}", flush=True)
    assert rel.max() <= TOL_C
    # per configuration the log|det| stays within FP32 evolution error over 128 slices
    assert np.abs(la - fp64_reference[0]).max() < 1e-5


def test_config3_xchain_only_within_1e6_of_fp64(ctx, fp64_reference):
    """ACC_XCHAIN keeps the FP64 x-chain but runs the y/z rows in FP32: about 8e-7 at the
    largest t at this shape, still inside 1e-6 at every t (and the z rows' rounding is the
    whole difference from ACC_X64, so it must stay well below the plain-FP32 figure)."""
    import nrcps_amd as na

    print("[gpu] FP32 (x-chain in FP64, z rows in FP32)", flush=True)
    la, ph = _fp32_run(ctx, True, na.ACC_XCHAIN)
    m64, c64 = _corr(*fp64_reference)
    m32, c32 = _corr(la, ph)
    rel = np.abs(c32 * np.exp(m32 - m64) - c64) / np.abs(c64)
    print(f"[gpu]   max rel err of C(t) {rel.max():.2e} at t={rel.argmax()}, median "
          f"{np.median(rel):.2e}", flush=True)
    assert rel.max() <= TOL_C
    assert np.abs(la - fp64_reference[0]).max() < 1e-5


def test_config3_plain_fp32_error_is_bounded(ctx, fp64_reference):
    """Every transform and multiply in FP32 (ACC_FP32) misses 1e-6 at some t at this shape
    (about 2e-6 at large t); this guards that figure so that a regression shows."""
    import nrcps_amd as na

    print("[gpu] FP32 (plain)", flush=True)
    la, ph = _fp32_run(ctx, True, na.ACC_FP32)
    m64, c64 = _corr(*fp64_reference)
    m32, c32 = _corr(la, ph)
    rel = np.abs(c32 * np.exp(m32 - m64) - c64) / np.abs(c64)
    print(f"[gpu]   max rel err of C(t) {rel.max():.2e}", flush=True)
    assert rel.max() <= 4e-6
    assert np.median(rel) <= 1e-6


@pytest.mark.parametrize("conj", [False, True])
def test_config3_dense_sinks_on_mfma_within_1e6_of_fp64(ctx, fp64_reference, conj):
    """The precision form as written: the overlaps as a dense GEMM on the matrix cores
    (split-BF16x6, FP32 accumulation) instead of the plane-wave gather, FP32 evolution with the
    FP64 x-chain four slices per launch (the states inside a launch written for the GEMM),
    FP64 LU; every orbital evolved, or conjugate partners (the stored volumes projected on the
    sinks and on their conjugate mirrors)."""
    import nrcps_amd as na

    print(f"[gpu] FP32 with dense plane-wave sinks (BF16x6 GEMM), partners={conj}", flush=True)
    moms = _lowest(L, N)
    V = L ** 3
    sinks = np.zeros((N, V), np.complex128)
    sinks[np.arange(N), (moms[:, 0] * L + moms[:, 1]) * L + moms[:, 2]] = 1.0
    e = na.Ensemble(ctx, L, moms, _xi_half(L), 0.4, NCONF, prec=na.PREC_F32, sinks=sinks)
    e.set_symmetry(conj)
    la, ph, _ = e.run(SEED, 0, T)
    e.close()
    m64, c64 = _corr(*fp64_reference)
    m32, c32 = _corr(la, ph)
    rel = np.abs(c32 * np.exp(m32 - m64) - c64) / np.abs(c64)
    print(f"[gpu]   max rel err of C(t) {rel.max():.2e} at t={rel.argmax()}, median "
          f"{np.median(rel):.2e}", flush=True)
    assert rel.max() <= TOL_C


def test_free_gas_t128_known_answer(ctx):
    """A = 0: plane waves are eigenvectors, det M(t) = prod_n K(1/2)(k_n)^(2t) exactly
    (log|det| = -2t sum log xi_half(k_n)); FP32 with the FP64 x-chain, conjugate partners."""
    import nrcps_amd as na

    C = 16
    moms, kin = _lowest(L, N), _xi_half(L)
    e = na.Ensemble(ctx, L, moms, kin, 0.0, C, prec=na.PREC_F32)
    e.set_symmetry(True)
    la, ph, _ = e.run(1, 0, T)
    e.close()
    idx = (moms[:, 0] * L + moms[:, 1]) * L + moms[:, 2]
    exact = -2 * np.arange(T)[:, None] * np.log(kin[idx]).sum()
    err = np.abs(la[:, :, 0] - exact).max()
    print(f"[gpu] free gas log|det| max abs error over T=128: {err:.2e}", flush=True)
    # Free plane waves are the coherent worst case: a wave with k_y != 0 is constant along x
    # and z, so the FP32 y-transform rounds the same value at every site and its error does
    # not average out as it does under a random field; it reaches ~9e-6 at t = 127 (the
    # random-field correlator above is held to 1e-6).
    assert err < 1.2e-5
    assert np.abs(ph[:, :, 0] - 1).max() < 1e-6
    # t = 0 measures the sources; their FP32 amplitudes divide out to rounding
    assert np.abs(la[0]).max() < 1e-12
