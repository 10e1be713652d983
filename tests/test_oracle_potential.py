"""Pins the oracle's Potential (the harmonic trap of src/hamiltonian/hamiltonian.C:163-240 and
Potential::Evolve, src/hamiltonian/noarch/hamiltonian_evolve.C:69-93) and the trapped slice
order K, P, I, P, K (production/slater_det3_trapped/main-cum.C:97-106) against closed forms
and a dense-matrix evolution at L = 4."""
import itertools

import numpy as np
import pytest


def _fold(n, L):
    return n - L if n >= L // 2 else n


@pytest.mark.parametrize("ptype", [0, 1])
def test_potential_table_closed_form(orc, ptype):
    L, k, dt = 8, (0.3, 0.05, 0.011), 0.5
    V = L ** 3
    got = orc.potential_table(L, orc.POT_FORM_HARMONIC, ptype, k, dt)
    for x, y, z in [(0, 0, 0), (5, 1, 7), (4, 4, 4), (3, 6, 2), (7, 7, 0)]:
        fx, fy, fz = _fold(x, L), _fold(y, L), _fold(z, L)
        v = (k[0] * fx * fx + k[1] * fy * fy + k[2] * fz * fz) / 2
        ref = (1 - dt * v if ptype == orc.POT_TYPE_LIN else np.exp(-dt * v)) / V
        assert abs(got[(x * L + y) * L + z] - ref) <= 4e-16 * abs(ref)  # numpy vs libm exp
    # no trap: 1/V everywhere; the fold puts -L/2 at n = L/2 only
    none = orc.potential_table(L, orc.POT_FORM_NONE, ptype, k, dt)
    assert np.all(none == 1.0 / V)
    t = got.reshape(L, L, L)
    m = (-np.arange(L)) % L
    inner = np.arange(1, L)  # away from n = L/2, where -L/2 has no mirror image inside
    inner = inner[inner != L // 2]
    sub = t[np.ix_(inner, inner, inner)]
    mir = t[np.ix_(m[inner], m[inner], m[inner])]
    assert np.array_equal(sub, mir)


def _dense_slice(L, kin, phi, A, pot):
    """The slice operator on k-space storage as a V x V matrix, from the brute-force DFT:
    K, then (if pot) F+ P F-, then F+ D F-, then (if pot) F+ P F-, then K."""
    V = L ** 3
    pts = np.array(list(itertools.product(range(L), repeat=3)))
    Fm = np.exp(-2j * np.pi * (pts @ pts.T) / L)  # sign -1 (FFTW_FORWARD)
    Fp = np.exp(+2j * np.pi * (pts @ pts.T) / L)
    K = np.diag(1.0 / kin)
    D = np.diag((1 + A * phi) / V)
    S = K
    if pot is not None:
        S = Fp @ np.diag(pot) @ Fm @ S
    S = Fp @ D @ Fm @ S
    if pot is not None:
        S = Fp @ np.diag(pot) @ Fm @ S
    return K @ S


def test_trapped_slices_match_dense_evolution(orc):
    L, N, T, A = 4, 3, 3, 0.4
    V = L ** 3
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 20.0, 0.5)
    pot = orc.potential_table(L, orc.POT_FORM_HARMONIC, orc.POT_TYPE_EXP, (0.9, 0.4, 0.2), 0.5)
    rng = np.random.default_rng(3)
    phi = rng.choice(np.array([-1, 1], np.int8), size=(T, V))
    got = orc.run_config(L, N, T, moms, kin, A=A, field="array", phi=phi, pot=pot,
                         want_psi=True, want_M=True)
    psi = np.zeros((N, V), complex)
    psi[np.arange(N), (moms[:, 0] * L + moms[:, 1]) * L + moms[:, 2]] = 1
    # the fused form the device uses: one x-space multiply by (V p)^2 (1 + A phi) / V
    Fm = np.exp(-2j * np.pi * (np.array(list(itertools.product(range(L), repeat=3))) @
                               np.array(list(itertools.product(range(L), repeat=3))).T) / L)
    Fp = Fm.conj()
    q = (V * pot) ** 2
    fused = psi.copy()
    for t in range(T):
        S = _dense_slice(L, kin, phi[t].astype(float), A, pot)
        psi = psi @ S.T
        Kd = 1.0 / kin
        fused = (((fused * Kd) @ Fm.T) * (q * (1 + A * phi[t]) / V)) @ Fp.T * Kd
    assert np.abs(got["psi"] - psi).max() <= 1e-12 * np.abs(psi).max()
    assert np.abs(fused - psi).max() <= 1e-12 * np.abs(psi).max()
    # the trap changes the result (a dropped P would be caught)
    free = orc.run_config(L, N, T, moms, kin, A=A, field="array", phi=phi, want_psi=True)
    assert np.abs(free["psi"] - got["psi"]).max() > 1e-3 * np.abs(psi).max()


def test_no_trap_potential_is_identity(orc):
    """POTENTIAL_FORM_NONE gives 1/V per site: each P is F+ (1/V) F- = identity."""
    L, N, T = 6, 7, 4
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 20.0, 0.5)
    pot = orc.potential_table(L, orc.POT_FORM_NONE, orc.POT_TYPE_EXP, (1, 1, 1), 0.5)
    a = orc.run_config(L, N, T, moms, kin, A=0.4, seed=9, config=2, pot=pot)
    b = orc.run_config(L, N, T, moms, kin, A=0.4, seed=9, config=2)
    assert np.abs(a["logabs"] - b["logabs"]).max() < 1e-12
