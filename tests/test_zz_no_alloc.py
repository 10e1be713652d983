"""nrc_overlap_noalloc (the overlap with its workspace in the device's fixed scratch array)
against nrc_overlap through the C ABI: the same kernels in the same order, so the results are
bit-identical; two calls in a row, the second smaller; sizes beyond the scratch refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import nrcps_amd as na

    c = na.Context(0)
    yield c
    c.close()


def _call(ctx, fn, phi_d, psi_d, M, ns, nt, V, ncfg, mode):
    return fn(ctx.h, phi_d.ptr, psi_d.ptr, M.ptr, ns, nt, V, ncfg, mode)


@pytest.mark.parametrize("mode", [3, 2, 0])  # BF16X6, BF16X3, F64
def test_noalloc_overlap_equals_overlap_twice(ctx, mode):
    import nrcps_amd as na

    rng = np.random.default_rng(7 + mode)
    dt = np.complex128 if mode == 0 else np.complex64
    for ns, nt, V, ncfg in ((33, 33, 4096, 8), (7, 5, 1728, 2)):  # the second call smaller
        print(f"[gpu] noalloc overlap mode={mode} ns={ns} nt={nt} V={V} ncfg={ncfg}", flush=True)
        phi = (rng.standard_normal((ns, V)) + 1j * rng.standard_normal((ns, V))).astype(dt)
        psi = (rng.standard_normal((ncfg, nt, V)) + 1j * rng.standard_normal((ncfg, nt, V))).astype(dt)
        phi_d, psi_d = ctx.to_device(phi), ctx.to_device(psi)
        M_ref = ctx.array((ncfg, ns, nt), np.complex128)
        M_na = ctx.array((ncfg, ns, nt), np.complex128)
        assert _call(ctx, ctx.lib.nrc_overlap, phi_d, psi_d, M_ref, ns, nt, V, ncfg, mode) == 0
        assert _call(ctx, ctx.lib.nrc_overlap_noalloc, phi_d, psi_d, M_na, ns, nt, V, ncfg,
                     mode) == 0
        ctx.sync()
        assert np.array_equal(M_na.download(), M_ref.download())


def test_noalloc_overlap_refuses_sizes_beyond_scratch(ctx):
    """ns = nt = 33 at V = 32768 for 1024 configurations needs ~0.3 GB of partial tiles: refused
    before any launch with NRC_ERR_TOO_LARGE."""
    import nrcps_amd as na

    phi_d = ctx.to_device(np.zeros(16, np.complex64))
    M = ctx.array((1,), np.complex128)
    rc = ctx.lib.nrc_overlap_noalloc(ctx.h, phi_d.ptr, phi_d.ptr, M.ptr, 33, 33, 32768, 1024, 3)
    assert rc == -7
    assert b"scratch" in ctx.lib.nrc_error_string(rc)
