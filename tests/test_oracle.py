"""Pins the FP64 CPU oracle against the reference's known answers, published vectors,
closed forms and brute force.  Every oracle function is covered here; a mistake such as a
dropped term, a wrong sign or index or a transposed operand fails at least one test."""
import ctypes as C
import itertools
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# ------------------------------------------------------------------------------ FFT
def _dft_matrix(n, sign):
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / n)


@pytest.mark.parametrize("L", [2, 3, 4, 6])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft3d_matches_bruteforce_dft(orc, L, sign):
    rng = np.random.default_rng(L)
    a = rng.standard_normal((L, L, L)) + 1j * rng.standard_normal((L, L, L))
    # full V x V DFT over the flattened site index z + L(y + L x)
    pts = np.array(list(itertools.product(range(L), repeat=3)))
    W = np.exp(sign * 2j * np.pi * (pts @ pts.T) / L)
    ref = (W @ a.reshape(-1)).reshape(L, L, L)
    got = orc.fft3d(a, sign)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("L", [12, 16, 24, 32])
def test_fft3d_matches_numpy(orc, L):
    rng = np.random.default_rng(100 + L)
    a = rng.standard_normal((3, L, L, L)) + 1j * rng.standard_normal((3, L, L, L))
    fwd = orc.fft3d(a, -1)  # FFTW_FORWARD == numpy fftn
    bwd = orc.fft3d(a, +1)  # FFTW_BACKWARD == V * numpy ifftn
    np.testing.assert_allclose(fwd, np.fft.fftn(a, axes=(1, 2, 3)), rtol=0, atol=1e-12 * L ** 3)
    np.testing.assert_allclose(bwd, np.fft.ifftn(a, axes=(1, 2, 3)) * L ** 3, rtol=0,
                               atol=1e-12 * L ** 3)


@pytest.mark.parametrize("shape", [(4, 6, 8), (2, 4, 64), (48, 8, 64), (32, 48, 12),
                                   (6, 64, 2), (64, 64, 64)])
def test_fft3d_non_cubic_matches_numpy(orc, shape):
    """X x Y x Z lattices and sides 48, 64: the axes are transformed with their own lengths
    and strides (a transposed stride or length fails this)."""
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal((1,) + shape) + 1j * rng.standard_normal((1,) + shape)
    V = int(np.prod(shape))
    np.testing.assert_allclose(orc.fft3d(a, -1), np.fft.fftn(a, axes=(1, 2, 3)), rtol=0,
                               atol=1e-12 * V)
    np.testing.assert_allclose(orc.fft3d(a, +1), np.fft.ifftn(a, axes=(1, 2, 3)) * V, rtol=0,
                               atol=1e-12 * V)


def test_fft3d_plane_wave_is_delta_and_roundtrip(orc):
    L = 8
    x = np.arange(L)
    k = (1, 3, 7)
    pw = np.exp(2j * np.pi * (k[0] * x[:, None, None] + k[1] * x[None, :, None]
                              + k[2] * x[None, None, :]) / L)
    f = orc.fft3d(pw, -1)
    ref = np.zeros_like(pw)
    ref[k] = L ** 3
    assert np.abs(f - ref).max() < 1e-11
    rng = np.random.default_rng(0)
    a = rng.standard_normal((L, L, L)) + 1j * rng.standard_normal((L, L, L))
    rt = orc.fft3d(orc.fft3d(a, -1), +1)
    assert np.abs(rt - L ** 3 * a).max() < 1e-11
    # Parseval
    assert abs(np.sum(np.abs(orc.fft3d(a, -1)) ** 2) - L ** 3 * np.sum(np.abs(a) ** 2)) < 1e-8


def test_fft1d_odd_radices(orc):
    for n in [5, 7, 9, 15, 18, 20]:
        a = np.random.default_rng(n).standard_normal((2, n)) * (1 + 0.5j)
        np.testing.assert_allclose(orc.fft1d(a, -1), np.fft.fft(a, axis=1), atol=1e-12)


# ------------------------------------------------------------------------------ tables
def _sites(L):
    g = np.arange(L)
    return np.meshgrid(g, g, g, indexing="ij")


def test_dispersion_standard_closed_form(orc):
    L, m = 12, 1.7
    X, Y, Z = _sites(L)
    ref = 2 * (np.sin(np.pi * X / L) ** 2 + np.sin(np.pi * Y / L) ** 2
               + np.sin(np.pi * Z / L) ** 2) / m
    np.testing.assert_allclose(orc.dispersion(L, orc.DISP_STANDARD, m), ref.reshape(-1),
                               rtol=1e-15, atol=1e-15)
    # antiperiodic in x shifts n_x by 1/2 (global_job_parameter.C:87-111)
    ref_a = 2 * (np.sin(np.pi * (X + 0.5) / L) ** 2 + np.sin(np.pi * Y / L) ** 2
                 + np.sin(np.pi * Z / L) ** 2) / m
    np.testing.assert_allclose(orc.dispersion(L, orc.DISP_STANDARD, m, bc=[1, 0, 0]),
                               ref_a.reshape(-1), rtol=1e-15, atol=1e-15)


def test_dispersion_quadratic_perfect_fold_and_cutoff(orc):
    L, m = 8, 1.18343
    X, Y, Z = _sites(L)
    f = lambda a: np.where(a >= L // 2, L - a, a)  # fold n >= L/2 to L - n (PRD)
    psq = 4 * np.pi ** 2 * ((f(X) / L) ** 2 + (f(Y) / L) ** 2 + (f(Z) / L) ** 2)
    q = orc.dispersion(L, orc.DISP_QUADRATIC, m, orc.CUTOFF_NONE)
    p = orc.dispersion(L, orc.DISP_PERFECT, m, orc.CUTOFF_NONE)
    np.testing.assert_allclose(q, (psq / (2 * m)).reshape(-1), rtol=1e-14)
    np.testing.assert_allclose(p, (np.expm1(psq / (2 * m))).reshape(-1), rtol=1e-13)
    c = orc.dispersion(L, orc.DISP_QUADRATIC, m, orc.CUTOFF_HARD, cutoff_sq=np.pi ** 2)
    beyond = (psq > np.pi ** 2).reshape(-1)
    assert np.all(np.isinf(c[beyond])) and np.all(np.isfinite(c[~beyond]))
    assert beyond.any() and (~beyond).any()


def test_kinetic_tables(orc):
    L, m, dt = 8, 2.0, 0.5
    xi = orc.kinetic_xi(L, orc.DISP_STANDARD, m, dt)
    np.testing.assert_allclose(xi, 1 + orc.dispersion(L, orc.DISP_STANDARD, m / dt), rtol=0)
    tau = 0.05
    k = orc.kinetic_exp(L, tau, 1.0)
    # exp(-tau eps) with eps = (1/2M) sum 4 sin^2(k/2) equals exp(-tau * STANDARD(mass=M))
    np.testing.assert_allclose(k, np.exp(-tau * orc.dispersion(L, orc.DISP_STANDARD, 1.0)),
                               rtol=1e-15)
    assert k[0] == 1.0 and np.all(k <= 1.0)


def test_interaction_single_coupling_closed_form(orc):
    # A = dt sqrt(4 pi c0 / M) (hamiltonian.C:120-136); A = 0.4 at M = 1, dt = 1
    L, A, M = 6, 0.4, 1.0
    c0 = A * A * M / (4 * np.pi)
    f, rc = orc.interaction_table(L, M, orc.INTER_ONEMINUSXIINVSQ, [c0], 1.0)
    assert rc == 0
    np.testing.assert_allclose(f, A / L ** 3, rtol=1e-14)
    phi = np.where(np.random.default_rng(3).random(L ** 3) < 0.5, -1.0, 1.0)
    fld = orc.interaction_field(phi.astype(complex), f, L)
    np.testing.assert_allclose(fld.real, (1 + A * phi) / L ** 3, rtol=0, atol=1e-16)
    assert np.abs(fld.imag).max() < 1e-17


def test_interaction_multi_coupling_matches_direct_formula(orc):
    # production/slater_det/args/interaction.arg
    L, M = 8, 5.0
    c = [9.82275, 2.06919, 0.45723, 0.00223528, 0.0451682, 0.00709158, 0.000244729]
    f, rc = orc.interaction_table(L, M, orc.INTER_ONEMINUSXIINVSQ, c, 1.0)
    assert rc == 0
    xi = np.minimum(1 + orc.dispersion(L, orc.DISP_PERFECT, M), np.exp(np.pi ** 2 / (2 * M)))
    psq = (1 - 1 / xi ** 2) * M
    ref = np.sqrt(sum(cj * psq ** j for j, cj in enumerate(c)) * 4 * np.pi / M) / L ** 3
    np.testing.assert_allclose(f, ref, rtol=1e-13)
    # the smeared field is real, has mean exactly (1 + f(0) V <phi>)/V
    phi = np.where(np.random.default_rng(4).random(L ** 3) < 0.5, -1.0, 1.0)
    fld = orc.interaction_field(phi.astype(complex), f, L)
    assert np.abs(fld.imag).max() < 1e-15
    assert abs(fld.real.mean() - (1 + f[0] * phi.sum()) / L ** 3) < 1e-15
    # it is the circular convolution of phi with the inverse transform of f
    ker = np.fft.ifftn(f.reshape(L, L, L)).real * L ** 3
    conv = np.fft.ifftn(np.fft.fftn(phi.reshape(L, L, L)) * np.fft.fftn(ker)).real
    np.testing.assert_allclose(fld.real, ((1 + conv) / L ** 3).reshape(-1), atol=1e-15)


# ------------------------------------------------------------------------------ momenta
def test_momenta_production_args(orc):
    # production/slater_det/args: L=16 PRD, PERFECT, E_F = 0.12, M = 1.18343
    n, moms, par, om = orc.momenta(16, orc.DISP_PERFECT, 1.18343, 0.12)
    assert n > 0
    assert np.all(np.diff(om[1:]) >= 0) and om[0] == 0.0
    # opposite-parity partners
    for i in range(n):
        assert np.array_equal((16 - moms[i]) % 16, moms[par[i]])
    # folded |n|^2 shells are complete and ordered
    f = np.where(moms > 8, moms - 16, moms)
    q = (f ** 2).sum(1)
    assert np.all(np.diff(q) >= 0)
    counts = np.bincount(q)
    full = np.bincount([a * a + b * b + c * c for a in range(-7, 9) for b in range(-7, 9)
                        for c in range(-7, 9)])
    assert np.array_equal(counts[counts > 0], full[: len(counts)][counts > 0])


@pytest.mark.parametrize("N,shell_end", [(1, 0), (7, 1), (19, 2), (27, 3), (33, 4)])
def test_lowest_momenta_closed_shells(orc, N, shell_end):
    L = 16
    m = orc.lowest_momenta(L, N)
    f = np.where(m > L // 2, m - L, m)
    q = (f ** 2).sum(1)
    assert q.max() == shell_end and np.all(np.diff(q) >= 0)
    assert len({tuple(r) for r in m}) == N
    # closed under k -> -k
    s = {tuple(r) for r in m}
    assert all(tuple((L - np.array(r)) % L) in s for r in s)


def test_lowest_momenta_partial_shell_tie_order(orc):
    # N = 14 at L = 24: 1 + 6 + the first 7 of the 12 |n|^2 = 2 states in enumeration order
    m = orc.lowest_momenta(24, 14)
    q2 = [tuple(r) for r in m[7:]]
    allq2 = [(a, b, c) for a in range(24) for b in range(24) for c in range(24)
             if sum((v - 24 if v > 12 else v) ** 2 for v in (a, b, c)) == 2]
    assert q2 == allq2[:7]


# ------------------------------------------------------------------------------ RNGs
def test_ranlxs_luscher_known_answer(orc):
    g = _load("ranlxs0_seed32767.json")
    r = orc.ranlxs(g["level"], g["seed"])
    xs = orc.ranlxs_draws(r, g["batches"] * g["batch_size"])
    last = xs[(g["batches"] - 1) * g["batch_size"]:]
    got = last[g["offset_in_last_batch"]: g["offset_in_last_batch"] + 96] * g["scale"]
    assert np.array_equal(got.astype(np.float64), np.array(g["values"]))


def test_ranlxs_state_roundtrip(orc):
    import ctypes as C

    r = orc.ranlxs(2, 7012)
    orc.ranlxs_draws(r, 1000)
    st = np.zeros(105, dtype=np.int64)
    orc.lib().orc_ranlxs_get_state(C.byref(r), st)
    a = orc.ranlxs_draws(r, 2040)
    r2 = orc.Ranlxs()
    assert orc.lib().orc_ranlxs_set_state(C.byref(r2), st) == 0
    b = orc.ranlxs_draws(r2, 2040)
    assert np.array_equal(a, b)
    assert st[0] == 105 and st[101] == 397
    bad = st.copy()
    bad[103] = (bad[102] + 6) % 12
    assert orc.lib().orc_ranlxs_set_state(C.byref(orc.Ranlxs()), bad) != 0


def test_ranlxs_z2_field_order(orc):
    # Lattice::Refresh ZTWO: phi_i = 2 floor(2 u_i) - 1 in site order
    import ctypes as C

    r1, r2 = orc.ranlxs(2, 7012), orc.ranlxs(2, 7012)
    u = orc.ranlxs_draws(r1, 500).astype(np.float64)
    phi = np.zeros(500, dtype=np.int8)
    orc.lib().orc_ranlxs_z2_field(C.byref(r2), 500, phi)
    assert np.array_equal(phi, (2 * np.floor(2 * u) - 1).astype(np.int8))


def test_philox_known_answers(orc):
    for v in _load("philox4x32_10_kat.json")["vectors"]:
        assert list(orc.philox(v["ctr"], v["key"])) == v["out"]


def test_z2_bits_layout_and_balance(orc):
    seed, V = 0x1234_5678_9ABC, 4096
    w = orc.z2_bits(seed, 5, 7, V)
    for g in range(len(w) // 4):
        out = orc.philox([g, 7, 5, 0], [seed & 0xFFFFFFFF, seed >> 32])
        assert np.array_equal(w[4 * g: 4 * g + 4], out)
    phi = orc.bits_to_phi(w, V)
    assert phi[0] == (1 if w[0] & 1 else -1) and phi[37] == (1 if (w[1] >> 5) & 1 else -1)
    assert abs(phi.astype(float).mean()) < 5 / np.sqrt(V)
    assert not np.array_equal(w, orc.z2_bits(seed, 5, 8, V))
    assert not np.array_equal(w, orc.z2_bits(seed, 6, 7, V))


# ------------------------------------------------------------------------------ LU
def test_ludcmp_reference_known_answer(orc):
    g = _load("ludcmp_known_answer.json")
    n = g["n"]
    a = np.array([[1.0 / (s + t + 2.0) for t in range(n)] for s in range(n)], dtype=complex)
    lu, indx, d = orc.ludcmp(a)
    la, ph, arg, det = orc.lu_det(lu, d)
    # the Cauchy determinant of 1/(x_s + y_t), x_s = s + 1, y_t = t + 1, in exact rationals
    x = [Fraction(s + 1) for s in range(n)]
    num = Fraction(1)
    den = Fraction(1)
    for i in range(n):
        for j in range(n):
            den *= x[i] + x[j]
            if i < j:
                num *= (x[j] - x[i]) ** 2
    exact = num / den
    assert abs(float(exact) / g["det"] - 1) < 1e-15
    assert abs(math.log(exact.numerator) - math.log(exact.denominator) - g["logdet"]) < 1e-12
    # FP64 LU of this cond ~1e14 matrix: agreement within the conditioning bound
    assert abs(det.real / g["det"] - 1) < 1e-2 and abs(la - g["logdet"]) < 1e-2
    assert ph == 1 and det.imag == 0
    rr = orc.ref_ludcmp(a)
    if rr is not None:
        assert np.array_equal(rr[0], lu) and np.array_equal(rr[1], indx) and rr[2] == d


@pytest.mark.parametrize("n", [2, 3, 10, 33, 64])
def test_ludcmp_oracle_equals_reference_build_on_gpu_test_inputs(orc, n):
    """The matrices of the GPU LU test (same seed, same tie cases): the oracle's factors,
    pivots and parity are the reference ludcmp.C's bit for bit, so the GPU test compares with
    the oracle alone and the reference build stays on this machine."""
    if not os.path.exists(orc.REF_LUDCMP_PATH):
        pytest.skip("oracle/_ref not built (no reference tree)")
    rng = np.random.default_rng(100 + n)
    Ms = rng.standard_normal((64, n, n)) + 1j * rng.standard_normal((64, n, n))
    if n == 10:
        Ms[0] = np.array([[1.0 / (s + t + 2.0) for t in range(n)] for s in range(n)])
        Ms[1] = np.eye(n)
    for c in range(64):
        lu, indx, d = orc.ludcmp(Ms[c])
        rr = orc.ref_ludcmp(Ms[c])
        assert np.array_equal(rr[0], lu) and np.array_equal(rr[1], indx) and rr[2] == d, c


def test_ludcmp_random_vs_lapack_and_reference_build(orc):
    rng = np.random.default_rng(11)
    for n in [1, 2, 3, 7, 19, 33, 48]:
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        lu, indx, d = orc.ludcmp(a)
        la, ph, arg, det = orc.lu_det(lu, d)
        sign, ld = np.linalg.slogdet(a)
        assert abs(la - ld) < 1e-12 * max(1, abs(ld))
        assert abs(ph - sign) < 1e-12
        assert abs(np.exp(1j * arg) - sign) < 1e-12
        # L U reproduces the row-permuted matrix
        P = np.arange(n)
        for j in range(n):
            P[[j, indx[j]]] = P[[indx[j], j]]
        Lm = np.tril(lu, -1) + np.eye(n)
        Um = np.triu(lu)
        np.testing.assert_allclose(Lm @ Um, a[P], atol=1e-12 * np.abs(a).max() * n)
        rr = orc.ref_ludcmp(a)
        if rr is not None:
            assert np.array_equal(rr[0], lu) and np.array_equal(rr[1], indx) and rr[2] == d


def test_ludcmp_tie_takes_last_row(orc):
    # equal scaled magnitudes in column 0: `>=` keeps the last maximal row (ludcmp.C:43)
    a = np.array([[1, 2, 0], [1, 0, 2], [1, 1, 1]], dtype=complex)
    _, indx, d = orc.ludcmp(a)
    rows = np.abs(a).max(1)
    scaled = np.abs(a[:, 0]) / rows
    assert indx[0] == np.flatnonzero(scaled == scaled.max())[-1]
    ident = np.eye(4, dtype=complex)
    _, indx, d = orc.ludcmp(ident)
    assert list(indx) == [0, 1, 2, 3] and d == 1.0


def test_ludcmp_singular_row_reported(orc):
    a = np.array([[1, 2], [0, 0]], dtype=complex)
    with pytest.raises(ZeroDivisionError):
        orc.ludcmp(a)


def test_overlap_is_conjugate_dot_and_unit_vector_gather(orc):
    rng = np.random.default_rng(5)
    phi = rng.standard_normal((3, 50)) + 1j * rng.standard_normal((3, 50))
    psi = rng.standard_normal((4, 50)) + 1j * rng.standard_normal((4, 50))
    np.testing.assert_allclose(orc.overlap(phi, psi), phi.conj() @ psi.T, rtol=1e-13)
    e = np.zeros((2, 50), complex)
    e[0, 7] = e[1, 31] = 1
    g = orc.overlap(e, psi)
    assert np.array_equal(g[0], psi[:, 7]) and np.array_equal(g[1], psi[:, 31])


# ------------------------------------------------------------------------------ the path
def _ref_tables(L, tau=0.05, M=1.0):
    """K-space split at tau=0.05, M=1: Kinetic(mass M/tau, dt=1/2), STANDARD."""
    return __import__("oracle").kinetic_xi(L, 0, M / tau, 0.5)


@pytest.mark.parametrize("split", [0, 1])
def test_free_gas_closed_form(orc, split):
    # A = 0: plane waves are eigenvectors, det M(t) = prod_n kappa(k_n)^t (cf. tests/free_gas)
    L, N, T = 8, 7, 6
    moms = orc.lowest_momenta(L, N)
    if split == orc.SPLIT_KSPACE:
        kin = _ref_tables(L)
        kap = 1 / kin ** 2
    else:
        kin = orc.kinetic_exp(L, 0.05, 1.0)
        kap = kin
    idx = (moms[:, 0] * L + moms[:, 1]) * L + moms[:, 2]
    r = orc.run_config(L, N, T, moms, kin, A=0.0, split=split, minors=True, want_M=True)
    for t in range(T):
        Mt = r["M"][t]
        np.testing.assert_allclose(Mt, np.diag(kap[idx] ** t), atol=1e-14)
        for n in range(1, N + 1):
            ref = t * np.log(kap[idx[:n]]).sum()
            assert abs(r["logabs"][t, n - 1] - ref) < 1e-12
            assert abs(r["phase"][t, n - 1] - 1) < 1e-12


def test_gather_timing_variant_equals_dense_overlaps(orc):
    """The CPU baseline's gather variant reads <k_s|psi_t> at k_s instead of the N^2
    full-length products: the same overlaps (the products add exact zeros), so the same
    determinants, at every slice."""
    L, N, T = 8, 7, 6
    moms = orc.lowest_momenta(L, N)
    kin = _ref_tables(L)
    a = orc.run_config(L, N, T, moms, kin, A=0.4, seed=3, config=2, minors=True, want_M=True)
    b = orc.run_config(L, N, T, moms, kin, A=0.4, seed=3, config=2, minors=True, want_M=True,
                       gather=True)
    for k in ("M", "logabs", "phase", "argsum"):
        assert np.array_equal(a[k], b[k]), k


def test_reference_split_smeared_field_equals_closed_form(orc):
    L, N, T, A = 6, 7, 5, 0.4
    moms = orc.lowest_momenta(L, N)
    kin = _ref_tables(L)
    f, _ = orc.interaction_table(L, 1.0, orc.INTER_ONEMINUSXIINVSQ, [A * A / (4 * np.pi)], 1.0)
    a = orc.run_config(L, N, T, moms, kin, A=A, seed=9, config=3)
    b = orc.run_config(L, N, T, moms, kin, A=A, seed=9, config=3, smear=f)
    np.testing.assert_allclose(a["logabs"], b["logabs"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(a["phase"], b["phase"], atol=1e-12)


def test_det_is_real_for_closed_shells(orc):
    L, N, T = 8, 7, 8
    moms = orc.lowest_momenta(L, N)
    r = orc.run_config(L, N, T, moms, _ref_tables(L), A=0.4, seed=1, config=0)
    assert np.all(np.abs(r["phase"].imag) < 1e-9)


def _s2(r):
    return np.exp(2 * r["logabs"][:, 0]) * (r["phase"][:, 0] ** 2).real


@pytest.mark.parametrize("split", [0, 1])
def test_exact_c1_matches_bruteforce_L2(orc, split):
    # every Z2 history of the two fields that reach the t = 2 measurement at L = 2 (V = 8)
    L, T, A = 2, 3, 0.4
    V = L ** 3
    moms = np.zeros((1, 3), dtype=np.int32)
    if split == orc.SPLIT_KSPACE:
        kin = _ref_tables(L, tau=0.3)
        ka = kb = 1 / kin
    else:
        kin = orc.kinetic_exp(L, 0.3, 1.0)
        ka, kb = kin, np.ones(V)
    exact = orc.exact_c1(L, T, ka, kb, A)
    acc = np.zeros(T)
    bits = np.arange(2 ** (2 * V))
    for h in bits:
        phi = np.array([1 if (h >> i) & 1 else -1 for i in range(2 * V)] + [1] * V, dtype=np.int8)
        r = orc.run_config(L, 1, T, moms, kin, A=A, split=split, field="array", phi=phi)
        acc += _s2(r)
    acc /= len(bits)
    np.testing.assert_allclose(acc, exact, rtol=1e-12)
    assert exact[0] == 1.0 and exact[1] > 1.0


def test_exact_c1_vs_monte_carlo_config1(orc):
    # L=12, T=32, tau=0.05, M=1, A=0.4, k=0 orbital, 500 configurations (x-space split:
    # x-space storage, psi <- IFFT[e^{-tau eps} FFT psi] then (1 + A phi))
    L, T, A, ncfg = 12, 32, 0.4, 500
    kin = orc.kinetic_exp(L, 0.05, 1.0)
    exact = orc.exact_c1(L, T, kin, np.ones(L ** 3), A)
    moms = np.zeros((1, 3), dtype=np.int32)
    s2 = np.array([_s2(orc.run_config(L, 1, T, moms, kin, A=A, split=orc.SPLIT_XSPACE,
                                      seed=7012, config=c)) for c in range(ncfg)])
    mean, err = s2.mean(0), s2.std(0, ddof=1) / np.sqrt(ncfg)
    assert np.all(np.abs(mean[1:] - exact[1:]) < 4 * err[1:] + 1e-12)
    assert mean[0] == pytest.approx(1.0, abs=1e-12)
    # E[S(t)] = 1 (E[1 + A phi] = 1 and K preserves the k = 0 mode)
    s = np.array([np.exp(orc.run_config(L, 1, 8, moms, kin, A=A, split=1, seed=5,
                                        config=c)["logabs"][:, 0]) for c in range(100)])
    assert abs(s[:, -1].mean() - 1) < 4 * s[:, -1].std() / 10


def test_gauss_field_is_standard_normal_and_shift_is_exact(orc):
    """Random::Gauss's polar Box-Muller (src/random/random.C:105-116): Kolmogorov-Smirnov
    against N(0, 1) (dropping the r2 division or taking x for y would fail it), acceptance
    rate pi/4 (2 uniforms per attempt: 8/pi uniforms per value), and Refresh(shift) adds the
    shift to the same draws."""
    from scipy import stats

    r = orc.ranlxs(2, 7012)
    g = orc.gauss_field(r, 200000, 0.0)
    assert stats.kstest(g, "norm").pvalue > 1e-3
    r1, r2 = orc.ranlxs(2, 99), orc.ranlxs(2, 99)
    assert np.array_equal(orc.gauss_field(r1, 1000, 0.0022),
                          orc.gauss_field(r2, 1000, 0.0) + 0.0022)
    # uniforms consumed: count draws until the two streams realign
    a, b = orc.ranlxs(2, 5), orc.ranlxs(2, 5)
    n_vals, n_unif = 20000, 0
    orc.gauss_field(a, n_vals, 0.0)
    sa = np.zeros(105, np.int64)
    orc.lib().orc_ranlxs_get_state(C.byref(a), sa)
    sb = np.zeros(105, np.int64)
    while True:
        orc.lib().orc_ranlxs_next(C.byref(b))
        n_unif += 1
        if n_unif % 2 == 0:
            orc.lib().orc_ranlxs_get_state(C.byref(b), sb)
            if np.array_equal(sa, sb):
                break
    assert abs(n_unif / n_vals - 8 / np.pi) < 0.03


def test_run_config_real_field_sinks_and_norms(orc):
    """The REAL field path equals the +-1 array path on +-1 values (with and without the
    smeared interaction); sinks 2 x sources add N log 2 to every log|det|; dividing by
    norms[t] after slice t subtracts N sum_{t' < t} log norms[t']."""
    L, N, T = 6, 7, 5
    V = L ** 3
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 20.0, 0.5)
    rng = np.random.default_rng(3)
    phi = np.where(rng.random((T, V)) < 0.5, 1, -1).astype(np.int8)
    f, rc = orc.interaction_table(L, 5.0, orc.INTER_ONEMINUSXIINVSQ, [0.5, 0.3], 1.0)
    for smear in (None, f):
        a = orc.run_config(L, N, T, moms, kin, smear=smear, field="array", phi=phi)
        b = orc.run_config(L, N, T, moms, kin, smear=smear, field="real",
                           phi_real=phi.astype(np.float64))
        assert np.array_equal(a["logabs"], b["logabs"])
    src = np.zeros((N, V), np.complex128)
    src[np.arange(N), (moms[:, 0] * L + moms[:, 1]) * L + moms[:, 2]] = 1.0
    base = orc.run_config(L, N, T, moms, kin, field="array", phi=phi)
    two = orc.run_config(L, N, T, moms, kin, field="array", phi=phi, sources=src,
                         sinks=2 * src)
    np.testing.assert_allclose(two["logabs"], base["logabs"] + N * np.log(2), atol=1e-12)
    norms = np.array([1.5, 0.7, 2.0, 1.1, 0.9])
    nrm = orc.run_config(L, N, T, moms, kin, field="array", phi=phi, norms=norms)
    shift = -N * np.concatenate([[0.0], np.cumsum(np.log(norms))[:-1]])
    np.testing.assert_allclose(nrm["logabs"][:, 0], base["logabs"][:, 0] + shift, atol=1e-12)
