"""Batches of uneven shapes (nrc_batch_*) and the autotuner (nrc_ensemble_tune), through the
C ABI: a batch must give every shape group exactly what the group gives when it runs alone
(the same kernels on the same inputs), with and without the captured HIP graph; the FP64
groups are also checked against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(8, 7, 3, 0), (12, 1, 5, 0), (16, 7, 2, 0), (8, 7, 4, 1), (12, 7, 2, 1)]  # L, N, C, prec


@pytest.fixture(scope="module")
def ctx():
    import nrcps_amd as na

    c = na.Context(0)
    yield c
    c.close()


def _groups(ctx, orc, shapes, split=0):
    import nrcps_amd as na

    out = []
    for L, N, C, prec in shapes:
        moms = orc.lowest_momenta(L, N)
        kin = (orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5) if split == 0
               else orc.kinetic_exp(L, 0.05, 1.0))
        out.append(na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=prec, split=split))
    return out


@pytest.mark.parametrize("graph", [1, 0])
def test_batch_matches_groups_alone(ctx, orc, graph):
    import nrcps_amd as na

    print(f"[gpu] batch of {len(SHAPES)} shape groups, graph={graph}", flush=True)
    T, seed = 9, 2024
    c0 = [10 * i + 1 for i in range(len(SHAPES))]
    alone = [g.run(seed, c, T) for g, c in zip(_groups(ctx, orc, SHAPES), c0)]
    groups = _groups(ctx, orc, SHAPES)
    b = na.Batch(ctx, groups)
    for rep in range(2):  # the second run replays the kept graph
        b.run(seed, c0, T, graph=bool(graph))
        for g, ref in zip(groups, alone):
            la, ph, ag = g.results(T)
            assert np.array_equal(la, ref[0]) and np.array_equal(ph, ref[1])
            assert np.array_equal(ag, ref[2])
    # new configurations and a new seed: the kept graph reads them from the parameter blocks
    c1 = [c + 100 for c in c0]
    for sd in (seed, seed + 1):
        b.run(sd, c1, T, graph=bool(graph))
        for g, (L, N, C, prec), c in zip(groups, SHAPES, c1):
            la, ph, _ = g.results(T)
            alone = _groups(ctx, orc, [(L, N, C, prec)])[0].run(sd, c, T)
            assert np.array_equal(la, alone[0]) and np.array_equal(ph, alone[1])
    assert b.captures() == (1 if graph else 0)
    b.run(seed, c1, T + 1, graph=bool(graph))  # another T: another graph
    assert b.captures() == (2 if graph else 0)
    b.close()
    # after the batch the ensembles run alone again
    la, _, _ = groups[0].run(seed, c0[0], T)
    assert np.array_equal(la, _groups(ctx, orc, SHAPES[:1])[0].run(seed, c0[0], T)[0])


def test_batch_fp64_groups_match_oracle(ctx, orc):
    import nrcps_amd as na

    shapes = [s for s in SHAPES if s[3] == 1]
    T, seed = 6, 77
    groups = _groups(ctx, orc, shapes)
    b = na.Batch(ctx, groups)
    b.run(seed, [5, 9], T, graph=True)
    for g, (L, N, C, _), c0 in zip(groups, shapes, [5, 9]):
        la, ph, _ = g.results(T)
        moms = orc.lowest_momenta(L, N)
        kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
        for c in range(C):
            ref = orc.run_config(L, N, T, moms, kin, A=0.4, split=orc.SPLIT_KSPACE, seed=seed,
                                 config=c0 + c)
            np.testing.assert_allclose(la[:, c, 0], ref["logabs"][:, 0], rtol=0, atol=1e-10)
            np.testing.assert_allclose(ph[:, c, 0], ref["phase"][:, 0], rtol=0, atol=1e-10)
    b.close()


def test_batch_rejects_shared_or_foreign_groups(ctx, orc):
    import nrcps_amd as na

    g = _groups(ctx, orc, SHAPES[:2])
    with pytest.raises(na.NrcError):
        na.Batch(ctx, [g[0], g[0]])
    b = na.Batch(ctx, g)
    with pytest.raises(na.NrcError):  # already in a batch
        na.Batch(ctx, [g[1]])
    b.close()


@pytest.mark.parametrize("L,prec", [(16, 0), (32, 0), (12, 1)])
def test_tune_picks_a_path_that_keeps_results(ctx, orc, L, prec):
    import nrcps_amd as na

    print(f"[gpu] tune L={L} prec={prec}", flush=True)
    N, C, T = 7, 3, 5
    g = _groups(ctx, orc, [(L, N, C, prec)])[0]
    path, ms = g.tune(3)
    print(f"[gpu] tuned L={L}: path {path}, {ms:.3f} ms/slice", flush=True)
    assert path in (na.PATH_TWO_PASS, na.PATH_SINGLE_PASS) and ms > 0
    la, ph, _ = g.run(31, 0, T)
    ref = _groups(ctx, orc, [(L, N, C, prec)])[0]
    ref.set_path(na.PATH_TWO_PASS)
    la2, ph2, _ = ref.run(31, 0, T)
    tol = 1e-11 if prec == 1 else 2e-5
    assert np.abs(la - la2).max() < tol and np.abs(ph - ph2).max() < tol


def test_batch_conj_partners_match_oracle_and_alone(ctx, orc):
    """Shape groups with conjugate partners in a captured batch: each group gives what it gives
    alone, and its FP64 determinants are the oracle's (which evolves every orbital)."""
    import nrcps_amd as na

    shapes = [(8, 7, 3, 1), (12, 19, 2, 1), (16, 27, 2, 0), (8, 1, 2, 1)]
    T, seed, c0 = 7, 31, [3, 40, 7, 11]
    alone = []
    for g, c in zip(_groups(ctx, orc, shapes), c0):
        g.set_symmetry(True)
        alone.append(g.run(seed, c, T))
    groups = _groups(ctx, orc, shapes)
    for g in groups:
        g.set_symmetry(True)
    b = na.Batch(ctx, groups)
    b.run(seed, c0, T, graph=True)
    for g, ref, (L, N, C, prec), c_0 in zip(groups, alone, shapes, c0):
        la, ph, _ = g.results(T)
        assert np.array_equal(la, ref[0]) and np.array_equal(ph, ref[1])
        if prec != 1:
            continue
        moms = orc.lowest_momenta(L, N)
        kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
        for c in range(C):
            o = orc.run_config(L, N, T, moms, kin, A=0.4, seed=seed, config=c_0 + c)
            np.testing.assert_allclose(la[:, c, 0], o["logabs"][:, 0], rtol=0, atol=1e-11)
            np.testing.assert_allclose(ph[:, c, 0], o["phase"][:, 0], rtol=0, atol=1e-11)
    # set_symmetry is refused while the ensemble is in a batch
    with pytest.raises(na.NrcError):
        groups[0].set_symmetry(False)
    b.close()


def test_batched_group_refuses_reallocating_calls(ctx, orc):
    """While an ensemble belongs to a batch its captured kernels hold its buffers and path, so
    every call that could reallocate them, change the path or layout, or free the ensemble is
    refused; after the batch is destroyed the same calls work and the batch's graph is gone."""
    import nrcps_amd as na

    print("[gpu] refused calls on a batched group", flush=True)
    g = _groups(ctx, orc, [(8, 7, 2, 1), (12, 1, 2, 0)])
    T, seed = 4, 5
    b = na.Batch(ctx, g)
    b.run(seed, [0, 0], T, graph=True)
    words = np.zeros((2 * T, 2, 16), np.uint32)
    for call in (lambda: g[0].run(seed, 0, 2 * T),  # larger T: the outputs would be reallocated
                 lambda: g[0].run_fields(words, 2 * T),
                 lambda: g[0].tune(2),
                 lambda: g[0].set_path(na.PATH_TWO_PASS),
                 lambda: g[0].set_symmetry(True),
                 lambda: g[0].set_trap(np.ones(8 ** 3)),
                 lambda: g[0].close()):
        with pytest.raises(na.NrcError, match="batch"):
            call()
    b.run(seed, [0, 0], T, graph=True)  # the kept graph is still valid
    la, _, _ = g[0].results(T)
    b.close()
    la2, _, _ = g[0].run(seed, 0, T)
    assert np.array_equal(la, la2)
    g[0].set_path(na.PATH_TWO_PASS)
    g[0].close()
