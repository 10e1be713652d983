"""The oracle of the memory network (oracle/memnet.py), pinned without the GPU: the memory
recurrence is the fields-with-memory one, a network with only its constant term is the tilted
sampler, the logit follows the stated FP32 order (recomputed in FP64 it differs by rounding
only), the cumulative log ratio of the sampler's own draws is its own, and the analytic
gradient of sum log q agrees with central differences."""
import numpy as np

import oracle
from oracle import convnet, memnet


def test_memory_recurrence_is_the_fields_with_memory_one():
    rng = np.random.default_rng(0)
    T, vol = 40, 64
    phis = np.where(rng.random((T, vol)) < 0.5, 1, -1)
    for r in memnet.R:
        M = np.zeros(vol, np.int64)
        m = np.zeros(vol)
        rho = r / 2.0 ** 15
        for t in range(T):
            M = memnet.mem_next(M, r, phis[t])
            m = rho * m + phis[t]
            assert np.abs(M / 1024.0 - m).max() < 0.05  # fixed point, rounded each slice
            assert np.abs(M).max() < 2 ** 15


def test_constant_network_is_the_tilted_sampler():
    lam = 3e-3
    p = np.zeros(memnet.NP, np.float32)
    p[34] = np.float32(2 * lam * 1024)
    M = np.arange(-3000, 3000, 7)
    l, _, _ = memnet.logits(p, M, -M)
    assert np.all(l == l[0])
    T = convnet.threshold(l)[0]
    assert abs(T / 2.0 ** 32 - (1 + np.tanh(lam)) / 2) < 1e-7


def test_logit_order_and_rounding():
    rng = np.random.default_rng(2)
    p = rng.standard_normal(memnet.NP).astype(np.float32)
    M0 = rng.integers(-4000, 4000, 1000)
    M1 = rng.integers(-16000, 16000, 1000)
    l, pre, _ = memnet.logits(p, M0, M1)
    assert l.dtype == np.float32 and pre.dtype == np.float32
    d = memnet.unpack(p)
    s0, s1 = M0 / 1024.0, M1 / 1024.0
    ref = d["c"][0] * s0 + d["c"][1] * s1 + d["b2"]
    for k in range(memnet.H):
        ref = ref + d["W2"][k] * np.maximum(d["W1"][k, 0] * s0 + d["W1"][k, 1] * s1
                                            + d["b1"][k], 0)
    ref = ref * 2.0 ** -10
    assert np.abs(l - ref).max() < 1e-5 * np.abs(ref).max()


def test_sampler_ratio_is_the_log_ratio_of_its_draws():
    p = np.random.default_rng(4).standard_normal(memnet.NP).astype(np.float32) * 4
    phis, want = memnet.sample(p, 7, 11, 5, 200)
    assert np.array_equal(memnet.log_ratio(p, phis), want)
    assert want[0] == 0.0 and np.all(want[1:] != 0.0)


def test_gradient_against_central_differences():
    rng = np.random.default_rng(1)
    p = rng.standard_normal(memnet.NP).astype(np.float32)
    # In the first slices every memory takes a few values only (all zero at t = 0), so many
    # sites share a hidden unit's input; small input weights and biases of +-2 keep every
    # unit on one side of its ReLU kink (four on, four off) for the differences.
    p[:16] *= np.float32(0.2)
    p[16:20] = 2.0
    p[20:24] = -2.0
    T, vol = 5, 400
    phis = np.where(rng.random((T, vol)) < 0.5, 1, -1).astype(np.int8)
    G = rng.random(T)

    def F(pp):
        tot = 0.0
        M0 = np.zeros(vol, np.int64)
        M1 = np.zeros(vol, np.int64)
        for t in range(T):
            l, _, _ = memnet.logits(pp, M0, M1)
            thr = convnet.threshold(l)
            q = np.where(phis[t] > 0, thr, 2.0 ** 32 - thr) / 2.0 ** 32
            tot += G[t] * np.log(q).sum()
            M0 = memnet.mem_next(M0, memnet.R[0], phis[t])
            M1 = memnet.mem_next(M1, memnet.R[1], phis[t])
        return tot

    g = memnet.grad_log_q(p, phis, G)
    h = 0.01
    for j in range(memnet.NP):
        e = np.zeros(memnet.NP, np.float32)
        e[j] = h
        fd = (F(p + e) - F(p - e)) / (2 * h)
        assert abs(fd - g[j]) < 2e-3 * np.abs(g).max() + 2e-2 * abs(g[j]), j
