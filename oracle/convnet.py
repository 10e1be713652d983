"""FP64 restatement of the learned field proposal (the network of nrc_proposal_*): forward
pass, weighted negative log-likelihood, its gradient, and the exact sampling ratio.

TEST INFRASTRUCTURE ONLY: the checker of the HIP proposal kernels; only tests/ import it.
The proposal has no counterpart in the reference (its fields are uniform Z2, src/alg/
lattice.C:81-86), so this module follows the definition in include/nrcps.h:
  taps 9 (dx+1) + 3 (dy+1) + (dz+1), out[s] = sum_tap W[tap] in[s + d(tap)], periodic;
  A1 = relu(conv(W1, [phi(t-1), u]) + b1), A2 = relu(conv(W2, A1) + b2),
  logit = conv(W3, A2) + b3, q(+1) = sigmoid(logit).
With bf16=True the weights and the stored activations (A1, A2 and the backward dZ1, dZ2) are
rounded to bfloat16 where the kernels store them, everything else in FP64.
Pinned in tests/test_oracle.py by a direct-loop convolution, finite differences of the loss
and the closed form of a bias-only network.
"""
from __future__ import annotations

import numpy as np

TAPS = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray) -> dict:
    p = np.asarray(p, dtype=np.float64)
    o = {}
    off = 0
    for name, shape in (("W1", (16, 27, 2)), ("b1", (16,)), ("W2", (16, 27, 16)), ("b2", (16,)),
                        ("W3", (27, 16)), ("b3", (1,))):
        n = int(np.prod(shape))
        o[name] = p[off:off + n].reshape(shape)
        off += n
    assert off == p.size == 8241
    return o


def pack(d: dict) -> np.ndarray:
    return np.concatenate([d[k].reshape(-1) for k in ("W1", "b1", "W2", "b2", "W3", "b3")])


def shift(x: np.ndarray, d) -> np.ndarray:
    """y[s] = x[s + d] on the periodic lattice; x is [B, L, L, L, ...]."""
    return np.roll(x, shift=(-d[0], -d[1], -d[2]), axis=(1, 2, 3))


def conv(x: np.ndarray, W: np.ndarray) -> np.ndarray:
    """x [B, L, L, L, ci], W [co, 27, ci] -> [B, L, L, L, co]."""
    out = 0.0
    for k, d in enumerate(TAPS):
        out = out + shift(x, d) @ W[:, k, :].T
    return out


def conv_t(y: np.ndarray, W: np.ndarray) -> np.ndarray:
    """The adjoint of conv: y [.., co] -> [.., ci], sum_tap W[co, tap, ci] y[s - d(tap)]."""
    out = 0.0
    for k, d in enumerate(TAPS):
        out = out + shift(y @ W[:, k, :], (-d[0], -d[1], -d[2]))
    return out


def bits_to_phi(words: np.ndarray, L: int) -> np.ndarray:
    """[B, nw] uint32 -> [B, L, L, L] of +-1 (bit j of word w is site 32 w + j)."""
    w = np.asarray(words, dtype=np.uint32)
    bits = ((w[:, :, None] >> np.arange(32, dtype=np.uint32)) & 1).reshape(w.shape[0], -1)
    return (2.0 * bits[:, :L ** 3] - 1.0).reshape(-1, L, L, L)


def forward(p, prev_phi, u, bf16_round=True):
    """prev_phi [B, L, L, L] (+-1) or None; u [B].  Returns logits [B, L, L, L] and the
    intermediates (z1, a1, z2, a2, x0)."""
    P = unpack(p)
    if bf16_round:
        P = {k: bf16(v) for k, v in P.items()}
    B = len(u)
    L = prev_phi.shape[1] if prev_phi is not None else None
    return _forward(P, prev_phi, np.asarray(u, dtype=np.float64), bf16_round, B, L)


def _forward(P, prev_phi, u, rnd, B, L):
    x0 = np.zeros((B, L, L, L, 2))
    if prev_phi is not None:
        x0[..., 0] = prev_phi
    x0[..., 1] = u[:, None, None, None]
    z1 = conv(x0, P["W1"]) + P["b1"]
    a1 = np.maximum(z1, 0.0)
    if rnd:
        a1 = bf16(a1)
    z2 = conv(a1, P["W2"]) + P["b2"]
    a2 = np.maximum(z2, 0.0)
    if rnd:
        a2 = bf16(a2)
    lg = conv(a2, P["W3"][None])[..., 0] + P["b3"][0]
    return lg, (z1, a1, z2, a2, x0)


def forward_L(p, prev_phi, u, L, bf16_round=True):
    """As forward, with the lattice side given (prev_phi may be None)."""
    P = unpack(p)
    if bf16_round:
        P = {k: bf16(v) for k, v in P.items()}
    return _forward(P, prev_phi, np.asarray(u, dtype=np.float64), bf16_round, len(u), L)


def loss_and_grad(p, prev_phi, target_phi, u, w, L, bf16_round=True):
    """loss = sum_b w_b sum_s -log q(target_b(s)) / (V sum_b w_b) and its gradient with
    respect to the parameter vector (the product path's definition)."""
    P = unpack(p)
    if bf16_round:
        P = {k: bf16(v) for k, v in P.items()}
    u = np.asarray(u, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    B, V = len(u), L ** 3
    lg, (z1, a1, z2, a2, x0) = _forward(P, prev_phi, u, bf16_round, B, L)
    y1 = (np.asarray(target_phi) > 0).astype(np.float64)
    a = np.where(y1 > 0, -lg, lg)
    nll = np.maximum(a, 0) + np.log1p(np.exp(-np.abs(a)))
    scale = 1.0 / (w.sum() * V)
    loss = scale * (w[:, None, None, None] * nll).sum()
    q = 1.0 / (1.0 + np.exp(-lg))
    dlg = scale * w[:, None, None, None] * (q - y1)  # [B, L, L, L]
    g = {}
    g["b3"] = np.array([dlg.sum()])
    g["W3"] = np.stack([np.einsum("bxyz,bxyzc->c", dlg, shift(a2, d)) for d in TAPS])
    da2 = conv_t(dlg[..., None], P["W3"][None])
    dz2 = da2 * (a2 > 0)
    if bf16_round:
        dz2 = bf16(dz2)
    g["b2"] = dz2.sum(axis=(0, 1, 2, 3))
    g["W2"] = np.stack([np.einsum("bxyzo,bxyzi->oi", dz2, shift(a1, d)) for d in TAPS], axis=1)
    da1 = conv_t(dz2, P["W2"])
    dz1 = da1 * (a1 > 0)
    if bf16_round:
        dz1 = bf16(dz1)
    g["b1"] = dz1.sum(axis=(0, 1, 2, 3))
    g["W1"] = np.stack([np.einsum("bxyzo,bxyzi->oi", dz1, shift(x0, d)) for d in TAPS], axis=1)
    return loss, pack(g), lg


def threshold(logit: np.ndarray) -> np.ndarray:
    """T with q(+1) = T / 2^32 (proposal_threshold, nrcps_amd/csrc/nrc_device.h): T =
    ceil(s 2^32) clamped to [1, 2^32 - 1], s = 1 / (1 + e), e = exp(-clamp(l, -30, 30)) by the
    Cody-Waite reduction and the degree-6 Taylor polynomial, every step one FP32 operation
    (numpy float32 arithmetic is IEEE, so this is bit-exact)."""
    f = np.float32
    l = np.asarray(logit, dtype=f)
    x = -np.minimum(np.maximum(l, f(-30.0)), f(30.0))
    n = np.rint(x * f(1.44269504))
    r = (x - n * f(0.693145751953125)) - n * f(1.42860677e-6)
    p = np.full_like(r, f(1.38888889e-3))
    for c in (8.33333333e-3, 4.16666667e-2, 0.166666667, 0.5, 1.0, 1.0):
        p = p * r + f(c)
    e = np.ldexp(p, n.astype(np.int32))
    s = f(1.0) / (f(1.0) + e)
    T = np.ceil(s * f(4294967296.0)).astype(np.float64)
    return np.clip(T, 1.0, 2.0 ** 32 - 1.0)


def sample_ratio(logit: np.ndarray, up: np.ndarray) -> np.ndarray:
    """Per-site log(0.5 / q(phi)) for the threshold draw: q(+1) = T / 2^32 with T =
    threshold(logit) (logit as the FP32 value); returns it and T."""
    T = threshold(logit)
    q = np.where(up, T, 2.0 ** 32 - T) / 2.0 ** 32
    return np.log(0.5 / q), T


def conv_direct(x: np.ndarray, W: np.ndarray) -> np.ndarray:
    """Direct loops over sites and taps (small lattices), for pinning `conv`."""
    B, L = x.shape[0], x.shape[1]
    out = np.zeros(x.shape[:4] + (W.shape[0],))
    for b in range(B):
        for i in range(L):
            for j in range(L):
                for k in range(L):
                    acc = np.zeros(W.shape[0])
                    for t, (dx, dy, dz) in enumerate(TAPS):
                        acc += W[:, t, :] @ x[b, (i + dx) % L, (j + dy) % L, (k + dz) % L]
                    out[b, i, j, k] = acc
    return out
