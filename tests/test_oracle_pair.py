"""The pair-sink functions of the oracle (TwoBody, SlaterDet3's measurement) pinned by
library values, closed forms and brute force: Dawson's integral against scipy and its ODE,
the opposite-parity index, the TwoBody wave functions, the pair overlap, and the free-gas
pair determinants of a whole run."""
import numpy as np
import pytest


def test_dawson_against_scipy_and_its_ode(orc):
    from scipy.special import dawsn

    x = np.concatenate([np.linspace(-6, 6, 1201), [0.0, 0.05, -0.19, 0.2, 0.21]])
    got = np.array([orc.dawson(v) for v in x])
    ref = dawsn(x)
    # Rybicki's sampling formula with step 0.4 (the reference's) is good to ~2e-7
    assert np.abs(got - ref).max() < 2e-7 * np.abs(ref).max()
    # below 0.2 the series stops at x^7: the first dropped term is 16 x^9 / 945
    small = np.abs(x) < 0.2
    assert np.all(np.abs(got[small] - ref[small]) <= 16 * np.abs(x[small]) ** 9 / 945 * 1.1
                  + 1e-16)
    # F'(x) = 1 - 2 x F(x)
    h = 1e-4
    for v in (0.3, 1.0, 2.5, 4.0):
        d = (orc.dawson(v + h) - orc.dawson(v - h)) / (2 * h)
        assert d == pytest.approx(1 - 2 * v * orc.dawson(v), abs=5e-6)
    assert orc.dawson(-1.3) == -orc.dawson(1.3)


@pytest.mark.parametrize("bc", [(0, 0, 0), (1, 0, 1)])
def test_opposite_parity_index(orc, bc):
    L = 6
    opp = orc.opposite_parity_index(L, bc).reshape(L, L, L)
    for x, y, z in [(0, 0, 0), (1, 2, 3), (5, 0, 4)]:
        q = opp[x, y, z]
        assert q == ((-(x + bc[0])) % L * L + (-(y + bc[1])) % L) * L + (-(z + bc[2])) % L
    if bc == (0, 0, 0):
        flat = opp.reshape(-1)
        assert np.array_equal(flat[flat], np.arange(L ** 3))  # an involution
        assert flat[0] == 0


def test_two_body_wavefunctions_closed_forms(orc):
    L, lam = 6, 0.7
    V = L ** 3
    n = np.arange(L)
    f = np.where(n > L // 2, n - L, n) if False else np.where(n >= (L + 1) // 2, L - n, n)
    p2 = (4 * np.pi ** 2 / L ** 2) * (f[:, None, None] ** 2 + f[None, :, None] ** 2
                                      + f[None, None, :] ** 2).reshape(-1)
    cut = np.pi ** 2
    assert np.all(orc.two_body_wavefunc(L, orc.WAVEFUNC_NONE, lam) == 0)
    assert np.allclose(orc.two_body_wavefunc(L, orc.WAVEFUNC_UNIFORM, lam), 1 / V)
    inside = p2 <= cut
    w1 = orc.two_body_wavefunc(L, orc.WAVEFUNC_PAIR1, lam, cutoff_sq=cut)
    np.testing.assert_allclose(w1[inside], 1 / (p2[inside] + lam), rtol=1e-13)
    assert np.all(w1[~inside] == 0)
    w2 = orc.two_body_wavefunc(L, orc.WAVEFUNC_PAIR2, lam, cutoff_sq=cut)
    sel = inside & (p2 > 0)
    np.testing.assert_allclose(w2[sel], np.exp(-lam * np.sqrt(p2[sel])) / p2[sel], rtol=1e-13)
    assert w2[0] == 0
    wt = orc.two_body_wavefunc(L, orc.WAVEFUNC_TGND, lam, cutoff_sq=cut)
    assert wt[0] == 1.0
    xs = np.sqrt(p2[sel]) / (2 * lam)
    np.testing.assert_allclose(wt[sel], [orc.dawson(v) / v for v in xs], rtol=1e-13)
    assert np.all(wt[~inside] == 0)
    # GND with the standard dispersion: xi = 1 + 2 sum sin^2(pi n / L) / m
    s = np.sin(np.pi * n / L) ** 2
    xi = 1 + 2 * (s[:, None, None] + s[None, :, None] + s[None, None, :]).reshape(-1) / 1.3
    wg = orc.two_body_wavefunc(L, orc.WAVEFUNC_GND, 3.0, orc.DISP_STANDARD, orc.DISP_STANDARD,
                               1.3, 1.3, cutoff_sq=cut)
    np.testing.assert_allclose(wg, 1 / (3.0 - 1 / (xi * xi)), rtol=1e-13)


def test_pair_overlap_brute_force_and_plane_waves(orc):
    rng = np.random.default_rng(2)
    L, n1, n2 = 4, 3, 5
    V = L ** 3
    a = rng.standard_normal((n1, V)) + 1j * rng.standard_normal((n1, V))
    b = rng.standard_normal((n2, V)) + 1j * rng.standard_normal((n2, V))
    w = rng.standard_normal(V)
    opp = orc.opposite_parity_index(L, (0, 1, 0))
    M = orc.pair_overlap(a, b, w, opp)
    ref = np.einsum("i,si,ti->st", w, a, b[:, opp])
    np.testing.assert_allclose(M, ref, rtol=1e-12, atol=1e-12)
    # unit vectors: M[s][t] = w(k_s) when k_t is the partner of k_s
    e = np.eye(V)[[0, 5, V - 5]].astype(np.complex128)
    Mp = orc.pair_overlap(e, e, w, orc.opposite_parity_index(L))
    opp0 = orc.opposite_parity_index(L)
    for s, ks in enumerate([0, 5, V - 5]):
        for t, kt in enumerate([0, 5, V - 5]):
            assert Mp[s, t] == (w[ks] if opp0[ks] == kt else 0.0)


def test_pair_run_free_gas_closed_form(orc):
    """A = 0: orbital n stays a plane wave with amplitude kappa_n^t (kappa = 1/xi_half^2), so
    the pair matrix is w(k_s) kappa_s^t kappa_t^t on partner pairs and the one-body overlaps
    are diagonal; closed shells pair +k with -k, so the leading pair minors of odd size and
    the replaced minors of even size are regular, the others exactly singular."""
    L, N, T = 6, 7, 3
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    w = orc.two_body_wavefunc(L, orc.WAVEFUNC_PAIR1, 0.5)
    opp = orc.opposite_parity_index(L)
    out = orc.run_config(L, N, T, moms, kin, A=0.0, pair_w=w, pair_opp=opp)
    idx = (moms[:, 0] * L + moms[:, 1]) * L + moms[:, 2]
    kap = 1 / kin[idx] ** 2
    for t in range(T):
        amp = kap ** t
        P = np.zeros((N, N), complex)
        for s in range(N):
            for q in range(N):
                if opp[idx[s]] == idx[q]:
                    P[s, q] = w[idx[s]] * amp[s] * amp[q]
        M1 = np.diag(amp).astype(complex)
        for n in range(N):
            for m, mat in ((2 * n + 1, P[:n + 1, :n + 1].copy()), (2 * n, None)):
                if mat is None:
                    mat = P[:n + 1, :n + 1].copy()
                    mat[:, n] = M1[n, :n + 1]
                sign, ld = np.linalg.slogdet(mat)
                got = out["logabs"][t, m]
                if np.isfinite(ld):
                    assert got == pytest.approx(ld, abs=1e-12), (t, m)
                    assert out["phase"][t, m] == pytest.approx(sign, abs=1e-12)
                else:  # zero in exact arithmetic, FFT rounding (~1e-17) in FP64
                    assert got < -25.0, (t, m)


def test_pair_run_minors_from_the_matrices(orc):
    """With a field, every determinant of the pair run equals numpy's on the matrices the
    same run exposes: leading minors of the pair matrix phi[s][t], and the odd case with
    column n taken from <sink_n | orbital_s> (slater_det3.C:163-167 with
    psi1[t + N s] = <sink_s | orbital_t>, slater_det3.C:97-104)."""
    L, N, T, seed = 4, 5, 3, 11
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    w = orc.two_body_wavefunc(L, orc.WAVEFUNC_PAIR1, 0.5)
    opp = orc.opposite_parity_index(L)
    one = orc.run_config(L, N, T, moms, kin, A=0.4, seed=seed, want_M=True)
    pair = orc.run_config(L, N, T, moms, kin, A=0.4, seed=seed, want_M=True, pair_w=w,
                          pair_opp=opp)
    for t in range(T):
        P, M1 = pair["M"][t], one["M"][t]
        if t > 0:  # the orientation of the column matters here
            assert not np.allclose(M1, M1.T)
        for n in range(N):
            a = P[:n + 1, :n + 1].copy()
            b = a.copy()
            b[:, n] = M1[n, :n + 1]
            for m, mat in ((2 * n + 1, a), (2 * n, b)):
                sign, ld = np.linalg.slogdet(mat)
                assert pair["logabs"][t, m] == pytest.approx(ld, abs=1e-11), (t, m)
                assert pair["phase"][t, m] == pytest.approx(sign, abs=1e-11), (t, m)
