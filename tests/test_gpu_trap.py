"""The trapped slice (nrc_ensemble_set_trap) against the oracle's literal reference order
K, P(1/2), I, P(1/2), K (production/slater_det3_trapped/main-cum.C:97-106), which applies each
Potential as its own forward / backward transform pair (src/hamiltonian/noarch/
hamiltonian_evolve.C:69-93); the device folds the three x-space multiplies into one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import nrcps_amd as na

    c = na.Context(0)
    yield c
    c.close()


def _trap(orc, L, ptype=1, k=(0.02, 0.02, 0.02), dt=0.5):
    pot = orc.potential_table(L, orc.POT_FORM_HARMONIC, ptype, k, dt)
    return pot, (L ** 3 * pot) ** 2


@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("ptype", [0, 1])
def test_trap_f64_vs_oracle(ctx, orc, path, ptype):
    import nrcps_amd as na

    L, N, T, C, seed = 12, 7, 10, 3, 41
    print(f"[gpu] trap f64 L={L} path={path} type={ptype}", flush=True)
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 20.0, 0.5)
    pot, q = _trap(orc, L, ptype)
    ens = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F64, minors=True)
    ens.set_path(path)
    ens.set_trap(q)
    la, ph, _ = ens.run(seed, 3, T)
    for c in range(C):
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, minors=True, seed=seed, config=3 + c,
                             pot=pot)
        assert np.abs(la[:, c] - ref["logabs"]).max() < 1e-12
        assert np.abs(ph[:, c] - ref["phase"]).max() < 1e-12
    # the trap changes the result; removing it gives the plain run back bit for bit
    ens.set_trap(None)
    la0, _, _ = ens.run(seed, 3, T)
    plain = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F64, minors=True)
    plain.set_path(path)
    assert np.array_equal(la0, plain.run(seed, 3, T)[0])
    assert np.abs(la0 - la).max() > 1e-4
    ens.close()
    plain.close()


def test_trap_f32_l32_with_partners(ctx, orc):
    """The trapped production shape (L=32, N=33 closed shells here) in FP32 with conjugate
    partners, against the oracle's FP64 literal order."""
    import nrcps_amd as na

    L, N, T, C, seed = 32, 33, 6, 2, 43
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 20.0, 0.5)
    pot, q = _trap(orc, L, 1, (0.000625, 0.000625, 0.000625))  # production potential.arg
    ens = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F32)
    ens.set_symmetry(True)
    ens.set_trap(q)
    la, ph, _ = ens.run(seed, 0, T)
    for c in range(C):
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, seed=seed, config=c, pot=pot)
        d = np.abs(la[:, c, 0] - ref["logabs"][:, 0]).max()
        print(f"[gpu] FP32 trapped, partners: max |d log|det|| {d:.2e}", flush=True)
        # the separable trap inside the FP64 x-chain kernel, four slices per launch (~2e-7)
        assert d < 1e-6
    ens.close()


def test_trap_with_smeared_field_and_partners(ctx, orc):
    import nrcps_amd as na

    L, N, T, C = 12, 7, 6, 2
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 20.0, 0.5)
    cpl = [9.82275, 2.06919, 0.45723, 0.00223528, 0.0451682, 0.00709158, 0.000244729]
    f, rc = orc.interaction_table(L, 5.0, orc.INTER_ONEMINUSXIINVSQ, cpl, 1.0)
    assert rc == 0
    f = f * 0.05
    pot, q = _trap(orc, L)
    ens = na.Ensemble(ctx, L, moms, kin, 0.0, C, prec=na.PREC_F64, smear=f)
    ens.set_symmetry(True)
    ens.set_trap(q)
    la, ph, _ = ens.run(5, 0, T)
    for c in range(C):
        ref = orc.run_config(L, N, T, moms, kin, smear=f, seed=5, config=c, pot=pot)
        assert np.abs(la[:, c] - ref["logabs"]).max() < 1e-11
        assert np.abs(ph[:, c] - ref["phase"]).max() < 1e-11
    ens.close()


def test_trap_refused_for_xspace_split(ctx, orc):
    import nrcps_amd as na

    L, N = 8, 7
    moms = orc.lowest_momenta(L, N)
    ens = na.Ensemble(ctx, L, moms, orc.kinetic_exp(L, 0.05, 1.0), 0.4, 2, prec=na.PREC_F64,
                      split=na.SPLIT_XSPACE)
    with pytest.raises(na.NrcError):
        ens.set_trap(np.ones(L ** 3))
    ens.close()


@pytest.mark.parametrize("L", [16, 24])
def test_trap_f32_four_slices_per_launch(ctx, orc, L):
    """FP32 trapped slices in the persistent kernel, four per launch (T = 9: launches of 4,
    4 and 1), against the oracle's literal order; the measurements between fused slices are
    gathered inside the kernel."""
    import nrcps_amd as na

    N, T, C, seed = 19, 9, 3, 47
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 20.0, 0.5)
    pot, q = _trap(orc, L, 1, (0.01, 0.02, 0.015))
    ens = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F32, minors=True)
    ens.set_trap(q)
    la, _, _ = ens.run(seed, 0, T)
    for c in range(C):
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, minors=True, seed=seed, config=c,
                             pot=pot)
        d = np.abs(la[:, c] - ref["logabs"]).max()
        print(f"[gpu] FP32 trapped L={L}, fused slices: max |d log|det|| {d:.2e}", flush=True)
        # FP64 x-chain; L = 16 with FP64 z-rows (~4e-7), L = 24 with FP32 ones (~4e-6)
        assert d < (2e-6 if L == 16 else 1e-5)
    ens.close()


@pytest.mark.parametrize("prec", [1, 0])
def test_trap_with_pair_sinks_and_partners(ctx, orc, prec):
    """SlaterDet3's measurement on trapped slices (the slater_det3_trapped loop) with conjugate
    partners, against the oracle's pair run with the literal K P I P K order: FP64 1e-9, FP32
    1e-4 + 1e-5 |log|det|| (the pair minors reach |log|det|| ~ 30 here)."""
    import nrcps_amd as na

    L, N, T, C, seed = 8, 7, 6, 2, 53
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    w = orc.two_body_wavefunc(L, orc.WAVEFUNC_TGND, 0.176776695, cutoff_sq=np.pi ** 2)
    opp = orc.opposite_parity_index(L)
    pot, q = _trap(orc, L, 1, (0.05, 0.04, 0.03))
    ens = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=prec, pair_w=w, pair_opp=opp)
    ens.set_symmetry(True)
    ens.set_trap(q)
    la, ph, _ = ens.run(seed, 0, T)
    for c in range(C):
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, seed=seed, config=c, pair_w=w,
                             pair_opp=opp, pot=pot)
        fin = np.isfinite(ref["logabs"])
        assert np.array_equal(np.isfinite(la[:, c, :]), fin)
        r = ref["logabs"][fin]
        tol = 1e-9 if prec == 1 else 1e-4 + 1e-5 * np.abs(r)
        assert np.all(np.abs(la[:, c, :][fin] - r) < tol)
    ens.close()
