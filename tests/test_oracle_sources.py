"""SHO sources of the oracle (OneBody::Set, SOURCE_TYPE_SHO): the Hermite recurrence pinned by
Gauss-Hermite orthonormality and closed forms, and the source of a single coefficient by its
Gaussian closed form (with lambda2 != lambda3, which pins the reference's scaling of the third
axis by lambda2)."""
import numpy as np


def test_hermite_orthonormal_and_closed_forms(orc):
    x, wq = np.polynomial.hermite.hermgauss(40)  # exact for polynomials of degree < 80
    H = np.array([orc.hermite(12, v) for v in x])  # [node][n]
    G = (H * wq[:, None]).T @ H
    np.testing.assert_allclose(G, np.eye(12), atol=1e-13)
    for v in (-1.3, 0.0, 0.7):
        h = orc.hermite(3, v)
        p = np.pi ** -0.25
        np.testing.assert_allclose(h, [p, np.sqrt(2) * v * p, (2 * v * v - 1) * p / np.sqrt(2)],
                                   rtol=1e-14, atol=1e-15)


def test_sho_source_single_coefficient_closed_form(orc):
    L, l1, l2, l3 = 8, 0.9, 1.3, 0.6
    c = np.zeros((10, 10, 10), np.complex128)
    c[0, 0, 0] = 0.5 - 0.25j
    c[1, 0, 2] = 0.0
    src = orc.sho_source(c, L, l1, l2, l3).reshape(L, L, L)
    k = np.arange(L)
    k = np.where(k >= L // 2, k - L, k)
    d1 = 2 * np.pi * k * l1 / L
    d2 = 2 * np.pi * k * l2 / L
    d3 = 2 * np.pi * k * l2 / L  # lambda2 on the third axis, as one_body.C:127
    g = np.exp(-(d1[:, None, None] ** 2 + d2[None, :, None] ** 2 + d3[None, None, :] ** 2) / 2)
    want = c[0, 0, 0] * np.pi ** -0.75 * g * np.sqrt(l1 * l2 * l3) * np.sqrt((2 * np.pi) ** 3 / L ** 3)
    np.testing.assert_allclose(src, want, rtol=1e-13, atol=1e-16)
    # a first-excited x state is odd in k_x (k and -k at x and L - x, x = 1 .. L/2 - 1)
    c[:] = 0
    c[1, 0, 0] = 1.0
    s1 = orc.sho_source(c, L, l1, l2, l3).reshape(L, L, L)
    x = np.arange(1, L // 2)
    np.testing.assert_allclose(s1[x], -s1[L - x], atol=1e-15)
    assert np.all(s1[0] == 0)
