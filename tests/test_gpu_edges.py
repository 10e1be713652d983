"""Edge cases at the boundary: empty batches, the smallest and largest sizes each entry point
takes and the first size it refuses, one configuration and one slice, odd run lengths
against the four-slice launches, and extreme tilts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import nrcps_amd as na

    c = na.Context(0)
    yield c
    c.close()


def _c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_empty_batches_are_no_ops(ctx):
    import nrcps_amd as na

    M = ctx.array((1, 2, 2), np.complex128)
    M.upload(np.full((1, 2, 2), 7 + 7j))
    phi = ctx.to_device(np.ones((2, 64), np.complex64))
    for mode in (na.OVERLAP_F32, na.OVERLAP_BF16X3, na.OVERLAP_BF16X6):
        ctx.overlap(phi, phi, M, 2, 2, 64, 0, mode)
    assert np.all(M.download() == 7 + 7j)  # nothing written
    d = ctx.to_device(np.ones((1, 4, 4, 4), np.complex64))
    ctx.fft3d(d, 4, 0, -1, na.PREC_F32)  # batch 0
    assert np.all(d.download() == 1)


def test_size_limits_are_refused_cleanly(ctx):
    import nrcps_amd as na

    big = ctx.array((1, 65, 65), np.complex128)
    la = np.zeros(1)
    with pytest.raises(na.NrcError):  # LU: n <= 64
        ctx.lu_det(big, 65, 1, False)
    phi = ctx.array((65, 64), np.complex64)
    M = ctx.array((1, 65, 65), np.complex128)
    with pytest.raises(na.NrcError):  # split-BF16 overlaps: ns, nt <= 64
        ctx.overlap(phi, phi, M, 65, 1, 64, 1, na.OVERLAP_BF16X3)
    with pytest.raises(na.NrcError):  # V must be a multiple of 4 for the MFMA forms
        ctx.overlap(phi, phi, M, 2, 2, 62, 1, na.OVERLAP_BF16X3)
    del la


def test_one_configuration_one_slice_and_odd_lengths(ctx, orc):
    """C = 1 on the L = 32 FP32 x-chain kernel (a grid far smaller than the CUs), T = 1 (the
    sources only) and T = 7 (four- and three-slice launches) against the FP64 path."""
    import nrcps_amd as na

    L, N = 32, 7
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    for T in (1, 7):
        f = na.Ensemble(ctx, L, moms, kin, 0.4, 1, prec=na.PREC_F32)
        f.set_symmetry(True)
        la32, ph32, _ = f.run(3, 11, T)
        f.close()
        r = na.Ensemble(ctx, L, moms, kin, 0.4, 1, prec=na.PREC_F64)
        la64, ph64, _ = r.run(3, 11, T)
        r.close()
        assert la32.shape == (T, 1, 1)
        assert np.abs(la32 - la64).max() < 2e-6 and np.abs(ph32 - ph64).max() < 2e-6
        assert abs(la64[0, 0, 0]) < 1e-12  # t = 0 measures the sources


@pytest.mark.parametrize("p", [2.0 ** -31, 1 - 2.0 ** -31])
def test_extreme_tilts_keep_exact_finite_ratios(ctx, orc, p):
    """q at one count from 0 or 1 (thr = 2 or 2^32 - 2): the fields are all one sign but a
    site in a billion, and the log ratios stay finite and equal to the oracle's for the
    drawn fields."""
    import nrcps_amd as na

    L, N, C, T = 8, 7, 2, 3
    V = L ** 3
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    e = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F64)
    q = e.set_tilt(p)
    thr = int(round(q * 2.0 ** 32))
    la, _, _ = e.run(5, 0, T)
    lr = e.log_ratio(T)
    e.close()
    assert np.all(np.isfinite(lr)) and np.all(np.isfinite(la))
    for c in range(C):
        phi = np.stack([orc.z2_tilt_phi(5, c, t, V, thr) for t in range(T)])
        assert np.allclose(lr[:, c], orc.z2_tilt_log_ratio(phi, q), rtol=1e-12)
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, field="array", phi=phi)
        assert np.abs(la[:, c] - ref["logabs"]).max() < 1e-10
