"""The autotuner beyond the path choice: slices per launch of the persistent kernels and the
overlap GEMM variants, the timings it reports, the choice it stores under the ensemble's
shape key, and the tuned choices saved to and loaded from a file by a fresh context."""
import os
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ens(ctx, orc, L, N, C, **kw):
    import nrcps_amd as na

    e = na.Ensemble(ctx, L, orc.lowest_momenta(L, N),
                    orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5), 0.4, C,
                    prec=na.PREC_F32, **kw)
    e.set_symmetry(True)
    return e


def test_tune_slices_per_launch_and_persist(orc):
    import nrcps_amd as na

    L, N, C, T = 16, 7, 64, 8
    print(f"[gpu] tune slices per launch L={L} N={N} C={C}", flush=True)
    ctx = na.Context(0)
    e = _ens(ctx, orc, L, N, C)
    ref, _, _ = e.run(5, 0, T)  # untuned: 4 slices per launch
    path, ms = e.tune(4)
    rep = e.tune_report()
    print(f"[gpu]   {rep}", flush=True)
    kinds = [k for k, _, _ in rep]
    assert kinds.count("path") == 2 and kinds.count("slices") == 3
    assert sorted(v for k, v, _ in rep if k == "slices") == [1, 2, 4]
    assert sorted(v for k, v, _ in rep if k == "tile") == [1, 2]  # x-pencils per thread
    assert all(t > 0 for _, _, t in rep) and ms > 0
    best = min((t, v) for k, v, t in rep if k == "slices")[1]
    assert e.launch_slices() == best
    la, _, _ = e.run(5, 0, T)
    assert np.abs(la - ref).max() < 2e-6  # launch boundaries move FP32 roundings only
    # a new ensemble of the same key takes the choice; another key does not
    f = _ens(ctx, orc, L, N, C)
    assert f.launch_slices() == best
    g = _ens(ctx, orc, L, N, C + 4)
    assert g.launch_slices() == 4
    with tempfile.TemporaryDirectory() as tmp:
        path_f = os.path.join(tmp, "tune.txt")
        ctx.tune_save(path_f)
        text = open(path_f).read()
        assert text.startswith("nrcps-tune 2\n") and "ensemble 16 7 0 0" in text
        ctx2 = na.Context(0)
        h = _ens(ctx2, orc, L, N, C)
        assert h.launch_slices() == 4
        ctx2.tune_load(path_f)
        assert h.launch_slices() == best
        bad = os.path.join(tmp, "bad.txt")
        with open(bad, "w") as fh:
            fh.write("nrcps-tune 2\nensemble 16 7 0 0 4 1 0 64 2 9 1\n")  # 9 slices: refused
        with pytest.raises(na.NrcError):
            ctx2.tune_load(bad)
        h.close()
        ctx2.close()
    for x in (e, f, g):
        x.close()
    ctx.close()


def test_tune_dense_sinks_picks_a_gemm_variant(orc):
    import nrcps_amd as na

    L, N, C, T = 16, 7, 32, 4
    V = L ** 3
    print(f"[gpu] tune dense sinks L={L} N={N} C={C}", flush=True)
    ctx = na.Context(0)
    moms = orc.lowest_momenta(L, N)
    sinks = np.zeros((N, V), np.complex128)
    sinks[np.arange(N), (moms[:, 0] * L + moms[:, 1]) * L + moms[:, 2]] = 1.0
    e = na.Ensemble(ctx, L, moms, orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5), 0.4, C,
                    prec=na.PREC_F32, sinks=sinks)
    ref, _, _ = e.run(9, 0, T)
    e.tune(2)
    rep = e.tune_report()
    print(f"[gpu]   {rep}", flush=True)
    gemm = {v: t for k, v, t in rep if k == "gemm"}
    assert sorted(gemm) == [na.OVL_COLS, na.OVL_SHARED, na.OVL_SHARED_DEEP, na.OVL_SHARED_PIPE]
    chosen = ctx.overlap_variant(N, N, V, na.OVERLAP_BF16X6)
    assert gemm[chosen] == min(gemm.values())
    la, _, _ = e.run(9, 0, T)
    assert np.abs(la - ref).max() < 1e-6  # the variants differ in FP32 summation order
    e.close()
    ctx.close()


def test_overlap_tune_and_pins(orc):
    """nrc_overlap_tune times every variant and keeps the fastest; a pinned variant is used
    until cleared; every variant gives the same overlaps to FP32 summation order."""
    import nrcps_amd as na

    ctx = na.Context(0)
    rng = np.random.default_rng(4)
    ns, nt, V, C = 19, 19, 2048, 8
    phi = (rng.standard_normal((ns, V)) + 1j * rng.standard_normal((ns, V))).astype(np.complex64)
    psi = (rng.standard_normal((C, nt, V)) + 1j * rng.standard_normal((C, nt, V))).astype(
        np.complex64)
    dphi, dpsi = ctx.to_device(phi), ctx.to_device(psi)
    M = ctx.array((C, ns, nt), np.complex128)
    mode = na.OVERLAP_BF16X6
    assert ctx.overlap_variant(ns, nt, V, mode) == na.OVL_SHARED  # the default
    best, ms = ctx.overlap_tune(dphi, dpsi, M, ns, nt, V, C, mode, reps=2)
    assert best == min(ms, key=ms.get) and ctx.overlap_variant(ns, nt, V, mode) == best
    outs = {}
    for v in (na.OVL_COLS, na.OVL_SHARED, na.OVL_SHARED_DEEP, na.OVL_SHARED_PIPE):
        ctx.set_overlap_variant(ns, nt, V, mode, v)
        assert ctx.overlap_variant(ns, nt, V, mode) == v
        ctx.overlap(dphi, dpsi, M, ns, nt, V, C, mode)
        outs[v] = M.download()
    for v, m in outs.items():
        assert np.abs(m - outs[na.OVL_COLS]).max() < 8e-8 * V
    ctx.set_overlap_variant(ns, nt, V, mode, na.OVL_AUTO)
    assert ctx.overlap_variant(ns, nt, V, mode) == na.OVL_SHARED
    with pytest.raises(na.NrcError):
        ctx.set_overlap_variant(ns, nt, V, mode, 9)
    ctx.close()
