"""Pins of the proposal network's FP64 restatement (oracle/convnet.py): the shifted-matmul
convolution against direct loops, the adjoint, the loss gradient against finite differences,
a bias-only network against its closed form, bf16 rounding and the sampling ratio."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def cn():
    from oracle import convnet

    return convnet


def test_conv_matches_direct_loops(cn):
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 4, 4, 4, 3))
    W = rng.standard_normal((5, 27, 3))
    np.testing.assert_allclose(cn.conv(x, W), cn.conv_direct(x, W), rtol=1e-12, atol=1e-12)
    # tap order: W nonzero only at tap (dx, dy, dz) = (1, 0, -1) picks x[s + d]
    W1 = np.zeros((1, 27, 1))
    W1[0, 9 * 2 + 3 * 1 + 0, 0] = 1.0
    y = cn.conv(x[..., :1], W1)
    assert np.array_equal(y[0, 1, 2, 3, 0], x[0, 2, 2, 2, 0])


def test_conv_t_is_the_adjoint(cn):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 4, 4, 3))
    y = rng.standard_normal((2, 4, 4, 4, 5))
    W = rng.standard_normal((5, 27, 3))
    lhs = (cn.conv(x, W) * y).sum()
    rhs = (x * cn.conv_t(y, W)).sum()
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_gradient_matches_finite_differences(cn):
    rng = np.random.default_rng(2)
    L, B = 4, 2
    p = rng.standard_normal(8241) * 0.3
    prev = np.where(rng.random((B, L, L, L)) < 0.5, 1.0, -1.0)
    tgt = np.where(rng.random((B, L, L, L)) < 0.6, 1.0, -1.0)
    u = rng.standard_normal(B)
    w = rng.random(B) + 0.5
    loss, g, _ = cn.loss_and_grad(p, prev, tgt, u, w, L, bf16_round=False)
    idx = np.concatenate([rng.choice(864, 6), 864 + rng.choice(16, 2), 880 + rng.choice(6912, 8),
                          7792 + rng.choice(16, 2), 7808 + rng.choice(432, 6), [8240]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (cn.loss_and_grad(pp, prev, tgt, u, w, L, bf16_round=False)[0]
              - cn.loss_and_grad(pm, prev, tgt, u, w, L, bf16_round=False)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_bias_only_network_closed_form(cn):
    L, B = 4, 3
    rng = np.random.default_rng(3)
    p = np.zeros(8241)
    b3 = 0.37
    p[8240] = b3
    tgt = np.where(rng.random((B, L, L, L)) < 0.7, 1.0, -1.0)
    w = np.array([1.0, 2.0, 0.5])
    loss, g, lg = cn.loss_and_grad(p, None, tgt, np.zeros(B), w, L, bf16_round=False)
    assert np.all(lg == b3)
    y = (tgt > 0).reshape(B, -1).mean(axis=1)
    q = 1 / (1 + np.exp(-b3))
    exp_loss = (w * -(y * np.log(q) + (1 - y) * np.log(1 - q))).sum() / w.sum()
    assert loss == pytest.approx(exp_loss, rel=1e-12)
    assert g[8240] == pytest.approx((w * (q - y)).sum() / w.sum(), rel=1e-12)
    assert np.all(g[:8240] == 0)


def test_bf16_rounding(cn):
    # 7 stored mantissa bits: the ulp of 1 is 2^-7
    x = np.array([1.0 + 2 ** -7, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5,
                  3e-39])
    r = cn.bf16(x)
    assert r[0] == 1.0 + 2 ** -7  # representable
    assert r[1] == 1.0  # tie, to even
    assert r[2] == 1.0 + 2 ** -6  # tie, to even
    assert r[3] == 1.0 + 2 ** -7  # nearest
    assert r[4] == -2.5 and r[5] > 0


def test_sample_ratio_is_normalised(cn):
    lg = np.array([-30.0, -1.0, 0.0, 1e-3, 0.7, 30.0], np.float32)
    r_up, T = cn.sample_ratio(lg, np.ones(6, bool))
    r_dn, _ = cn.sample_ratio(lg, np.zeros(6, bool))
    q_up, q_dn = 0.5 / np.exp(r_up), 0.5 / np.exp(r_dn)
    np.testing.assert_allclose(q_up + q_dn, 1.0, rtol=1e-15)
    assert T[2] == 2.0 ** 31 and r_up[2] == 0.0 and r_dn[2] == 0.0
    assert 1.0 <= T.min() and T.max() <= 2.0 ** 32 - 1


def test_threshold_tracks_the_sigmoid(cn):
    """The FP32 threshold recipe is monotone, exact at 0 and within 1e-7 of sigmoid * 2^32."""
    lg = np.linspace(-35.0, 35.0, 200001).astype(np.float32)
    T = cn.threshold(lg)
    sig = 1.0 / (1.0 + np.exp(-lg.astype(np.float64)))
    assert np.abs(T / 2.0 ** 32 - sig).max() < 1e-7
    assert np.all(np.diff(T) >= 0)
    assert cn.threshold(np.float32([0.0]))[0] == 2.0 ** 31
    assert T.min() == 1.0 and T.max() == 2.0 ** 32 - 1
    # a contraction of p * r + c into one fused multiply-add changes T somewhere: the recipe
    # is the one with separately rounded products
    f = np.float32
    x = -np.clip(lg, f(-30), f(30))
    n = np.rint(x * f(1.44269504))
    r = (x - n * f(0.693145751953125)) - n * f(1.42860677e-6)
    p = np.full_like(r, f(1.38888889e-3)).astype(np.float64)
    for c in (8.33333333e-3, 4.16666667e-2, 0.166666667, 0.5, 1.0, 1.0):
        p = (p * r.astype(np.float64) + np.float64(f(c))).astype(np.float32).astype(np.float64)
    e = np.ldexp(p.astype(np.float32), n.astype(np.int32))
    Tf = np.clip(np.ceil((f(1) / (f(1) + e)) * f(4294967296.0)).astype(np.float64), 1, 2 ** 32 - 1)
    assert np.any(Tf != T)
