"""Conjugate partners (nrc_ensemble_set_symmetry) against the FP64 oracle, which evolves every
orbital as the reference does (src/propagator/propagator.C:35-64).

With real fields and even kinetic tables the orbital of momentum -k is the conjugate mirror of
the one of momentum k, so evolving one of each pair must give the oracle's determinants:
1e-12 in FP64, the FP32 evolution's own error in FP32."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import nrcps_amd as na

    c = na.Context(0)
    yield c
    c.close()


def _tables(orc, L, split, tau=0.05):
    if split == 0:
        return orc.kinetic_xi(L, orc.DISP_STANDARD, 1.0 / tau, 0.5)
    return orc.kinetic_exp(L, tau, 1.0)


def _moms_partial(orc, L):
    """k = 0, +-e_x, +e_y and -e_z: two pairs, one self-conjugate and two orbitals without a
    partner in the set."""
    return np.array([[0, 0, 0], [1, 0, 0], [L - 1, 0, 0], [0, 1, 0], [0, 0, L - 1]], np.int32)


@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("shape", ["closed", "partial"])
def test_conj_partners_f64_vs_oracle(ctx, orc, split, path, shape):
    import nrcps_amd as na

    L, T, C, seed = 12, 24, 3, 99
    moms = orc.lowest_momenta(L, 19) if shape == "closed" else _moms_partial(orc, L)
    N = len(moms)
    print(f"[gpu] conj partners f64 L={L} N={N} split={split} path={path}", flush=True)
    kin = _tables(orc, L, split)
    ens = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F64, split=split, minors=True)
    ens.set_path(path)
    ens.set_symmetry(True)
    la, ph, _ = ens.run(seed, 7, T)
    for c in range(C):
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, split=split, minors=True, seed=seed,
                             config=7 + c)
        dla = np.abs(la[:, c] - ref["logabs"]).max()
        dph = np.abs(ph[:, c] - ref["phase"]).max()
        assert dla < 1e-12 and dph < 1e-12, (c, dla, dph)
    ens.close()


@pytest.mark.parametrize("split", [0, 1])
def test_conj_partners_orbitals_vs_oracle(ctx, orc, split):
    import nrcps_amd as na

    L, N, T = 8, 7, 5
    moms = orc.lowest_momenta(L, N)
    kin = _tables(orc, L, split)
    ens = na.Ensemble(ctx, L, moms, kin, 0.4, 2, prec=na.PREC_F64, split=split)
    with pytest.raises(na.NrcError):  # no run yet
        ens.orbitals(0)
    ens.run(3, 40, 2)
    ens.set_symmetry(True)
    with pytest.raises(na.NrcError):  # stored in the layout before the change
        ens.orbitals(0)
    ens.run(3, 40, T)
    ref = orc.run_config(L, N, T, moms, kin, A=0.4, split=split, seed=3, config=41,
                         want_psi=True)["psi"]
    got = ens.orbitals(1)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    ens.close()


def test_conj_partners_smeared_field(ctx, orc):
    """The smeared field of an even form factor is real: partners hold (hamiltonian.C:77-147)."""
    import nrcps_amd as na

    L, N, T, C = 12, 7, 8, 2
    moms = orc.lowest_momenta(L, N)
    kin = _tables(orc, L, 0)
    cpl = [9.82275, 2.06919, 0.45723, 0.00223528, 0.0451682, 0.00709158, 0.000244729]
    f, rc = orc.interaction_table(L, 5.0, orc.INTER_ONEMINUSXIINVSQ, cpl, 1.0)
    assert rc == 0
    f = f * 0.05
    ens = na.Ensemble(ctx, L, moms, kin, 0.0, C, prec=na.PREC_F64, smear=f)
    ens.set_symmetry(True)
    la, ph, _ = ens.run(5, 0, T)
    for c in range(C):
        ref = orc.run_config(L, N, T, moms, kin, smear=f, seed=5, config=c)
        assert np.abs(la[:, c] - ref["logabs"]).max() < 1e-11
        assert np.abs(ph[:, c] - ref["phase"]).max() < 1e-11
    ens.close()


def test_conj_partners_f32_l32_n33(ctx, orc):
    """The benchmark shape (L=32, N=33: 17 volumes per configuration) in FP32 with four slices
    per launch, against the oracle's FP64 determinants."""
    import nrcps_amd as na

    L, N, T, C = 32, 33, 12, 2
    moms = orc.lowest_momenta(L, N)
    kin = _tables(orc, L, 0)
    ens = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F32, minors=True)
    ens.set_symmetry(True)
    la, ph, _ = ens.run(21, 0, T)
    for c in range(C):
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, minors=True, seed=21, config=c)
        d = np.abs(la[:, c] - ref["logabs"]).max()
        print(f"[gpu] FP32 partners log|det| max abs err (config {c}): {d:.2e}", flush=True)
        assert d < 2e-6  # every leading minor; ~3e-7 measured with the FP64 x-chain
        # the full determinant is real: conjugate pairs make it so column by column
        assert np.abs(ph[:, c, N - 1].imag).max() < 1e-5
    ens.close()


def test_conj_partners_same_as_all_orbitals_f32(ctx, orc):
    """Partners on and off on the same fields: the same log|det| to FP32 evolution error."""
    import nrcps_amd as na

    L, N, T, C = 16, 27, 16, 8
    moms = orc.lowest_momenta(L, N)
    kin = _tables(orc, L, 0)
    a = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F32)
    la0, _, _ = a.run(5, 0, T)
    a.set_symmetry(True)
    la1, _, _ = a.run(5, 0, T)
    a.set_symmetry(False)
    la2, _, _ = a.run(5, 0, T)
    assert np.abs(la0 - la1).max() < 2e-6  # FP32 state, FP64 x-chain: ~1e-7 per run
    assert np.array_equal(la0, la2)  # off again: the all-orbital run, bit for bit
    a.close()


def test_conj_partners_rejected_where_not_exact(ctx, orc):
    import nrcps_amd as na

    L, N = 8, 7
    moms = orc.lowest_momenta(L, N)
    kin = _tables(orc, L, 0)
    V = L ** 3
    rng = np.random.default_rng(0)
    # a kinetic table that is not even under p -> -p
    odd = kin * (1 + 1e-3 * rng.random(V))
    e = na.Ensemble(ctx, L, moms, odd, 0.4, 2, prec=na.PREC_F64)
    with pytest.raises(na.NrcError):
        e.set_symmetry(True)
    e.close()


@pytest.mark.parametrize("prec,split", [(1, 0), (1, 1), (0, 0), (0, 1)])
def test_dense_sinks_with_conj_partners(ctx, orc, prec, split):
    """Dense sinks with conjugate partners: the sinks and their conjugate mirrors J phi(p) =
    conj phi(-p) are projected on the stored volumes and a partner's column is the conjugate
    of its mirrored-sink overlaps.  Random dense sinks (no symmetry of their own), against the
    oracle, which evolves every orbital and projects on the sinks as given: FP64 1e-11; FP32
    against its own every-orbital run (the same GEMM, separately rounded partner volumes)."""
    import nrcps_amd as na

    L, N, T, C, seed = 8, 7, 6, 3, 41
    print(f"[gpu] dense sinks with conjugate partners prec={prec} split={split}", flush=True)
    moms = orc.lowest_momenta(L, N)
    kin = _tables(orc, L, split)
    V = L ** 3
    rng = np.random.default_rng(17)
    sinks = rng.standard_normal((N, V)) + 1j * rng.standard_normal((N, V))
    sinks /= np.linalg.norm(sinks, axis=1, keepdims=True)
    e = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=prec, split=split, sinks=sinks)
    e.set_symmetry(True)
    assert e.volumes_per_config() == 4
    la, ph, _ = e.run(seed, 0, T)
    e.close()
    if prec == na.PREC_F64:
        for c in range(C):
            ref = orc.run_config(L, N, T, moms, kin, A=0.4, split=split, seed=seed, config=c,
                                 sinks=sinks)
            err = np.abs(la[:, c, 0] - ref["logabs"][:, 0]).max()
            perr = np.abs(ph[:, c, 0] - ref["phase"][:, 0]).max()
            print(f"[gpu] config {c}: max |d log|det|| {err:.2e}, phase {perr:.2e}", flush=True)
            assert err < 1e-11 and perr < 1e-10
    else:
        f = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=prec, split=split, sinks=sinks)
        lf, pf, _ = f.run(seed, 0, T)
        f.close()
        err, perr = np.abs(la - lf).max(), np.abs(ph - pf).max()
        print(f"[gpu] FP32 partners vs every orbital: {err:.2e}, phase {perr:.2e}", flush=True)
        assert err < 2e-5 and perr < 2e-5
