"""Oracle of the memory-network proposal.

TEST INFRASTRUCTURE ONLY: the checker of the HIP memory-network kernels; only tests/ import it.

No counterpart in the reference: its fields are uniform Z2 draws (src/alg/lattice.C:81-86);
the proposal is an importance sampler for them, exact like the shifted Gaussian of
production/one_body-reweight/main.C:73-76.

Every site carries two decayed memories of its own past field, M_j = 1024 m_j with m_j(t) =
sum_{k >= 1} rho_j^(k-1) phi(t - k) in fixed point: M_j(0) = 0, M_j(t + 1) =
floor((r_j M_j + 2^14) / 2^15) + 1024 phi(t), r_j = round(2^15 rho_j) (the recurrence of the
fields with memory).  A pointwise network maps s_j = M_j / 1024 to a logit

    a_k = max((W1[k][0] s_0 + W1[k][1] s_1) + b1[k], 0),        k = 0 .. 7
    l   = ((c_0 s_0 + c_1 s_1) + b2 + sum_k W2[k] a_k) * 2^-10

(every operation one FP32 rounding, in this order, no fused multiply-adds), and site i is +1
where its Philox output (counter (i / 4, t, config, 4), output i % 4) is below T(l), the
threshold of oracle/convnet.py: q(+1) = T / 2^32 exactly.  The slice's log(p / q) is
sum_i log(2^31 / f_i), f_i = T_i or 2^32 - T_i.

Parameters, 35 FP32 values: W1 [8][2], b1 [8], W2 [8], c [2], b2.
"""
import numpy as np

from . import convnet

H = 8
NP = 2 * H + H + H + 2 + 1
assert NP == 35
RHO = (0.75, 0.94)
R = tuple(int(round(r * 2 ** 15)) for r in RHO)
SCALE = np.float32(2.0 ** -10)
STREAM = 4


def unpack(p):
    p = np.asarray(p, dtype=np.float32)
    assert p.shape == (NP,)
    return {"W1": p[:16].reshape(H, 2), "b1": p[16:24], "W2": p[24:32], "c": p[32:34],
            "b2": p[34]}


def pack(d) -> np.ndarray:
    return np.concatenate([np.asarray(d["W1"], np.float32).reshape(-1),
                           np.asarray(d["b1"], np.float32), np.asarray(d["W2"], np.float32),
                           np.asarray(d["c"], np.float32),
                           np.asarray([d["b2"]], np.float32)]).astype(np.float32)


def mem_next(M: np.ndarray, r: int, phi: np.ndarray) -> np.ndarray:
    return ((r * M.astype(np.int64) + (1 << 14)) >> 15) + 1024 * phi.astype(np.int64)


def logits(p, M0: np.ndarray, M1: np.ndarray):
    """FP32 logits and the hidden pre-activations for memories M0, M1 (integer arrays)."""
    d = unpack(p)
    f = np.float32
    s0 = M0.astype(f) * f(1.0 / 1024)
    s1 = M1.astype(f) * f(1.0 / 1024)
    acc = d["c"][0] * s0
    acc = acc + d["c"][1] * s1
    acc = acc + d["b2"]
    pre = []
    for k in range(H):
        z = d["W1"][k, 0] * s0
        z = z + d["W1"][k, 1] * s1
        z = z + d["b1"][k]
        pre.append(z)
        acc = acc + d["W2"][k] * np.maximum(z, f(0.0))
    return acc * SCALE, np.stack(pre), (s0, s1)


def site_log_ratio(T: np.ndarray, up: np.ndarray) -> np.ndarray:
    f = np.where(up, T, 2.0 ** 32 - T)
    return np.log(2.0 ** 31 / f)


def sample(p, seed: int, config: int, T: int, vol: int):
    """The sampler: fields phis [T][vol] (int8) of one configuration and the cumulative
    log(p / q) before each slice, [T]."""
    from . import z2_bits

    M0 = np.zeros(vol, np.int64)
    M1 = np.zeros(vol, np.int64)
    phis = np.empty((T, vol), np.int8)
    inc = np.zeros(T)
    for t in range(T):
        u = z2_bits(seed, config, t, 32 * vol, stream=STREAM).astype(np.float64)
        l, _, _ = logits(p, M0, M1)
        thr = convnet.threshold(l)
        up = u < thr
        phi = np.where(up, 1, -1)
        inc[t] = site_log_ratio(thr, up).sum()
        M0, M1 = mem_next(M0, R[0], phi), mem_next(M1, R[1], phi)
        phis[t] = phi
    return phis, np.concatenate([[0.0], np.cumsum(inc)[:-1]])


def log_ratio(p, phis: np.ndarray) -> np.ndarray:
    """Cumulative log(p / q) [T] that the proposal gives fields phis [T][vol] (drawn by any
    sampler)."""
    T, vol = phis.shape
    M0 = np.zeros(vol, np.int64)
    M1 = np.zeros(vol, np.int64)
    inc = np.zeros(T)
    for t in range(T):
        l, _, _ = logits(p, M0, M1)
        up = phis[t] > 0
        inc[t] = site_log_ratio(convnet.threshold(l), up).sum()
        M0, M1 = mem_next(M0, R[0], phis[t]), mem_next(M1, R[1], phis[t])
    return np.concatenate([[0.0], np.cumsum(inc)[:-1]])


def grad_log_q(p, phis: np.ndarray, G: np.ndarray) -> np.ndarray:
    """sum_t G[t] sum_i d log q(phi_i(t)) / d params, with q(+1) = sigmoid(l) (the smooth
    form of T(l) / 2^32, which it matches to ~2e-7), in FP64 from the FP32 features."""
    T, vol = phis.shape
    d = unpack(p)
    M0 = np.zeros(vol, np.int64)
    M1 = np.zeros(vol, np.int64)
    g = np.zeros(NP)
    for t in range(T):
        l, pre, (s0, s1) = logits(p, M0, M1)
        up = phis[t] > 0
        sig = 1.0 / (1.0 + np.exp(-l.astype(np.float64)))
        dl = G[t] * (up.astype(np.float64) - sig) * float(SCALE)  # d log q / d (l / SCALE)
        s0d, s1d = s0.astype(np.float64), s1.astype(np.float64)
        act = pre.astype(np.float64) > 0
        gW1 = np.zeros((H, 2))
        gb1 = np.zeros(H)
        gW2 = np.zeros(H)
        for k in range(H):
            back = dl * act[k] * float(d["W2"][k])
            gW1[k, 0] = (back * s0d).sum()
            gW1[k, 1] = (back * s1d).sum()
            gb1[k] = back.sum()
            gW2[k] = (dl * np.maximum(pre[k].astype(np.float64), 0.0)).sum()
        g += np.concatenate([gW1.reshape(-1), gb1, gW2, [(dl * s0d).sum(), (dl * s1d).sum()],
                             [dl.sum()]])
        M0, M1 = mem_next(M0, R[0], phis[t]), mem_next(M1, R[1], phis[t])
    return g
