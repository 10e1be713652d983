"""The learned auxiliary-field proposal (nrc_proposal_* in libnrcps_hip.so) and the sampler and
trainer built on it.

The reference draws every field uniformly (src/alg/lattice.C:81-86).  Here a small 3D
convolutional network proposes phi(t) given phi(t-1) and the configuration's normalised
log-weight at t; every site's exact proposal probability enters the importance ratio
r_t = prod over slices < t and sites of (1/2) / q(phi), so that mean(det(t)^2 r_t) is an
unbiased estimate of the uniform-field correlator C(t) = <det(t)^2>.  The network is trained
by weighted maximum likelihood on fields of the uniform sampler, weighted by det^2.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import Context, DeviceArray, Ensemble, NrcError, _check, _ptr

STREAM_UNIFORM = 0  # Philox stream of the uniform sampler (nrc_ensemble_run)
STREAM_PROPOSAL = 1  # Philox stream of the proposal's site draws

N_HIDDEN, N_TAP = 16, 27


def param_layout():
    """(name, offset, shape) of the parameter vector."""
    out, off = [], 0
    for name, shape in (("W1", (16, 27, 2)), ("b1", (16,)), ("W2", (16, 27, 16)), ("b2", (16,)),
                        ("W3", (27, 16)), ("b3", (1,))):
        n = int(np.prod(shape))
        out.append((name, off, shape))
        off += n
    return out


def unpack(params: np.ndarray) -> dict:
    return {name: params[off:off + int(np.prod(shape))].reshape(shape)
            for name, off, shape in param_layout()}


def init_params(rng: np.random.Generator, out_scale: float = 1e-3) -> np.ndarray:
    """He-scaled hidden layers; the output layer starts small, so q starts near uniform."""
    p = np.zeros(8241, np.float32)
    v = unpack(p)
    v["W1"][:] = rng.standard_normal(v["W1"].shape) * np.sqrt(2.0 / (27 * 2))
    v["W2"][:] = rng.standard_normal(v["W2"].shape) * np.sqrt(2.0 / (27 * 16))
    v["W3"][:] = rng.standard_normal(v["W3"].shape) * out_scale
    v["b1"][:] = 0.1
    v["b2"][:] = 0.1
    return p


class Proposal:
    """nrc_proposal_*: parameters on the device, forward, sampling and training steps."""

    def __init__(self, ctx: Context, L: int, max_batch: int, params: np.ndarray | None = None):
        self.ctx, self.lib, self.L, self.max_batch = ctx, ctx.lib, L, max_batch
        self.V, self.nw = L ** 3, L ** 3 // 32
        self.n_params = self.lib.nrc_proposal_param_count()
        h = C.c_void_p()
        _check(self.lib.nrc_proposal_create(ctx.h, L, max_batch, C.byref(h)), "proposal_create")
        self.h = h
        self.set_params(init_params(np.random.default_rng(0)) if params is None else params)

    def set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.n_params
        _check(self.lib.nrc_proposal_set_params(self.h, _ptr(p)), "proposal_set_params")

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        _check(self.lib.nrc_proposal_get_params(self.h, _ptr(p)), "proposal_get_params")
        return p

    def forward(self, prev: DeviceArray | None, u: DeviceArray, B: int, logits: DeviceArray):
        _check(self.lib.nrc_proposal_forward(self.h, prev.ptr if prev else None, u.ptr, B,
                                             logits.ptr), "proposal_forward")

    def sample(self, prev, u, B, seed, config0, t, words: DeviceArray, logr: DeviceArray,
               stream=STREAM_PROPOSAL):
        _check(self.lib.nrc_proposal_sample(self.h, prev.ptr if prev else None, u.ptr, B, seed,
                                            config0, t, stream, words.ptr, logr.ptr),
               "proposal_sample")

    def train_step(self, prev, target, u, w, B, wsum, lr, want_grad=False, optimizer=0):
        """optimizer 0: Adam, 1: SGD with momentum 0.9."""
        g = np.zeros(self.n_params, np.float32) if want_grad else None
        loss = C.c_double()
        _check(self.lib.nrc_proposal_train_step(self.h, prev.ptr if prev else None, target.ptr,
                                                u.ptr, w.ptr, B, wsum, lr, optimizer,
                                                _ptr(g) if want_grad else None, C.byref(loss)),
               "proposal_train_step")
        return loss.value, g

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.lib.nrc_proposal_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def norm_logw(ctx: Context, logabs_dev: int, stride: int, B: int, mu: float, sd: float,
              u: DeviceArray):
    _check(ctx.lib.nrc_norm_logw(ctx.h, C.c_void_p(logabs_dev), stride, B, mu, sd, u.ptr),
           "norm_logw")


def z2_bits_list(ctx: Context, seed: int, configs: DeviceArray, ts: DeviceArray, n: int, vol: int,
                 words: DeviceArray, stream=STREAM_UNIFORM):
    _check(ctx.lib.nrc_z2_bits_list(ctx.h, seed, configs.ptr, ts.ptr, n, stream, vol, words.ptr),
           "z2_bits_list")


def logw_stats(la: np.ndarray, floor: float = 1e-3):
    """Per-slice mean and spread of 2 log|det| over configurations (la [T][C] or [T][C][1])."""
    x = 2.0 * (la[:, :, 0] if la.ndim == 3 else la)
    return x.mean(axis=1), np.maximum(x.std(axis=1), floor)


def det2_weights(la: np.ndarray, ts) -> np.ndarray:
    """Training weights of the uniform sample: the mean over the slices ts of det(t)^2 /
    sum_c det(t)^2."""
    x = la[:, :, 0] if la.ndim == 3 else la
    w = np.zeros(x.shape[1])
    for t in ts:
        e = np.exp(2.0 * (x[t] - x[t].max()))
        w += e / e.sum()
    return w / len(ts)


class _Slot:
    """A device address inside a larger allocation (one slice of a [S][B][nw] field array)."""

    def __init__(self, addr: int):
        self.ptr = C.c_void_p(addr)


class LearnedSampler:
    """Configurations sampled slice by slice from the proposal: measure at t, normalise the
    log-weight, draw phi(t), evolve slice t (nrc_ensemble_begin / measure / evolve)."""

    def __init__(self, ctx: Context, prop: Proposal, ens: Ensemble, mu: np.ndarray,
                 sd: np.ndarray):
        assert ens.C <= prop.max_batch and ens.L == prop.L
        self.ctx, self.prop, self.ens, self.mu, self.sd = ctx, prop, ens, mu, sd
        B = ens.C
        self.words = [ctx.array((B, prop.nw), np.uint32), ctx.array((B, prop.nw), np.uint32)]
        self.u = ctx.array((B,), np.float32)
        self.logr = ctx.array((B,), np.float64)
        self.group_words = None

    def run(self, seed: int, config0: int, T: int, group: int = 1):
        """Returns log|det| [T][C][m], phase [T][C][m] and log r_t [T][C] (the ratio of the
        slices before t).  group > 1: the fields of `group` slices are drawn in a row, each
        conditioned on the previous slice's field and on the log-weight of the last row
        measured before the group (0 before any), and evolved in one launch of the persistent
        kernels (nrc_ensemble_evolve_fused); the ratios stay exact for the fields drawn."""
        if group > 1:
            return self._run_grouped(seed, config0, T, group)
        e, B = self.ens, self.ens.C
        e.begin(T)
        la_dev, _ = e.results_dev()
        _check(self.ctx.lib.nrc_memset(self.ctx.h, self.logr.ptr, 0, self.logr.nbytes), "memset")
        logr_hist = self.ctx.array((T, B), np.float64)
        prev = None
        for t in range(T):
            e.measure(t)
            _check(self.ctx.lib.nrc_memcpy_d2d(self.ctx.h, C.c_void_p(logr_hist.ptr.value + t * B * 8),
                                               self.logr.ptr, B * 8), "d2d")
            norm_logw(self.ctx, la_dev + t * B * e.nm * 8, e.nm, B, float(self.mu[t]),
                      float(self.sd[t]), self.u)
            cur = self.words[t & 1]
            self.prop.sample(prev, self.u, B, seed, config0, t, cur, self.logr)
            e.evolve(t, cur)
            prev = cur
        la, ph, _ = e.results(T)
        return la, ph, logr_hist.download()

    def _run_grouped(self, seed: int, config0: int, T: int, group: int):
        e, B, nw = self.ens, self.ens.C, self.prop.nw
        if self.group_words is None or self.group_words.shape[0] < group:
            self.group_words = self.ctx.array((group, B, nw), np.uint32)
        gw = self.group_words
        e.begin(T)
        la_dev, _ = e.results_dev()
        lib, h = self.ctx.lib, self.ctx.h
        _check(lib.nrc_memset(h, self.logr.ptr, 0, self.logr.nbytes), "memset")
        logr_hist = self.ctx.array((T, B), np.float64)
        prev = None
        for t0 in range(0, T, group):
            S = min(group, T - t0)
            if t0 == 0:
                _check(lib.nrc_memset(h, self.u.ptr, 0, self.u.nbytes), "memset")
            else:
                norm_logw(self.ctx, la_dev + (t0 - 1) * B * e.nm * 8, e.nm, B,
                          float(self.mu[t0 - 1]), float(self.sd[t0 - 1]), self.u)
            for k in range(S):
                t = t0 + k
                _check(lib.nrc_memcpy_d2d(h, C.c_void_p(logr_hist.ptr.value + t * B * 8),
                                          self.logr.ptr, B * 8), "d2d")
                cur = _Slot(gw.ptr.value + k * B * nw * 4)
                self.prop.sample(prev, self.u, B, seed, config0, t, cur, self.logr)
                prev = cur
            e.evolve_fused(t0, S, gw)
        la, ph, _ = e.results(T)
        return la, ph, logr_hist.download()


def train(ctx: Context, prop: Proposal, la: np.ndarray, seed: int, config0: int, steps: int,
          batch: int, lr: float, ts_weight, rng: np.random.Generator, log_every: int = 0,
          optimizer: int = 0):
    """Weighted maximum likelihood on the uniform sampler's fields: configurations c are drawn
    in proportion to their det^2 weights (det2_weights over ts_weight), slices t uniformly;
    the sample is (phi_c(t-1), u_c(t)) -> phi_c(t), fields regenerated from Philox(seed).
    The learning rate decays to zero on a cosine; Adam (optimizer 0) by default, SGD with
    momentum (optimizer 1) on request.
    Returns the loss history."""
    T, Cn = la.shape[0], la.shape[1]
    mu, sd = logw_stats(la)
    w = det2_weights(la, ts_weight)
    x2 = 2.0 * (la[:, :, 0] if la.ndim == 3 else la)
    V = prop.V
    cfg_d = ctx.array((batch,), np.uint32)
    t_d = ctx.array((batch,), np.uint32)
    tm1_d = ctx.array((batch,), np.uint32)
    prev_d = ctx.array((batch, prop.nw), np.uint32)
    tgt_d = ctx.array((batch, prop.nw), np.uint32)
    u_d = ctx.array((batch,), np.float32)
    w_d = ctx.array((batch,), np.float32).upload(np.ones(batch, np.float32))
    losses = []
    for step in range(steps):
        c = rng.choice(Cn, size=batch, p=w)
        first = step % T == 0  # a batch of first slices (no previous field) now and then
        t = np.zeros(batch, np.int64) if first else rng.integers(1, T, size=batch)
        u = ((x2[t, c] - mu[t]) / sd[t]).astype(np.float32)
        cfg_d.upload((config0 + c).astype(np.uint32))
        t_d.upload(t.astype(np.uint32))
        z2_bits_list(ctx, seed, cfg_d, t_d, batch, V, tgt_d)
        if not first:
            tm1_d.upload((t - 1).astype(np.uint32))
            z2_bits_list(ctx, seed, cfg_d, tm1_d, batch, V, prev_d)
        u_d.upload(u)
        lr_t = 0.5 * lr * (1.0 + np.cos(np.pi * step / steps))  # cosine decay to zero
        loss, _ = prop.train_step(None if first else prev_d, tgt_d, u_d, w_d, batch,
                                  float(batch), lr_t, optimizer=optimizer)
        losses.append(loss)
        if log_every and step % log_every == 0:
            print(f"  train step {step}: loss {loss:.8f}", flush=True)
    return np.array(losses), mu, sd


def z2_sums(ctx: Context, seed: int, config0: int, nconf: int, T: int, vol: int,
            stream=STREAM_UNIFORM) -> np.ndarray:
    """S[t][c] = sum over sites of the plain sampler's phi(x, t) (nrc_z2_sums)."""
    out = ctx.array((T, nconf), np.int32)
    _check(ctx.lib.nrc_z2_sums(ctx.h, seed, config0, nconf, T, stream, vol, out.ptr), "z2_sums")
    return out.download()


def memory_log_ratio(ctx: Context, seed: int, config0: int, nconf: int, T: int, vol: int,
                     lam: float, theta: float, rho: float) -> np.ndarray:
    """[T][nconf] cumulative log(p / q) that fields with memory (lam, theta, rho) would carry,
    evaluated on the plain sampler's fields of configurations config0 .. config0 + nconf - 1
    (nrc_z2_memory_eval)."""
    out = ctx.array((T, nconf), np.float64)
    _check(ctx.lib.nrc_z2_memory_eval(ctx.h, seed, config0, nconf, T, vol, float(lam),
                                      float(theta), float(rho), out.ptr), "z2_memory_eval")
    return out.download()


def _lme(x: np.ndarray) -> float:
    """log of the mean of exp(x)."""
    m = x.max()
    return float(np.log(np.exp(x - m).mean()) + m)


def memory_log_ess(la: np.ndarray, lr: np.ndarray, ts) -> float:
    """Mean over t in ts of the log effective-sample fraction that a proposal q with
    log(p / q) = lr [T][C] (on a plain sample) gives the estimator of <det(t)^2>:
    E_p[d2]^2 / E_p[d2^2 p / q] (la [T][C][m] of the same configurations)."""
    x = la[:, :, 0] if la.ndim == 3 else la
    return float(np.mean([2 * _lme(2 * x[t]) - _lme(4 * x[t] + lr[t]) for t in ts]))


def fit_memory(ctx: Context, la: np.ndarray, seed: int, config0: int, T: int, vol: int, ts,
               lam0: float, theta0: float = 2.5e-4, rho0: float = 0.88, maxiter: int = 100,
               nfit: int = 0):
    """(lam, theta, rho) of fields with memory that maximise memory_log_ess on a plain sample
    (la of configurations config0 ..; their fields regenerated on the device), Nelder-Mead in
    scaled coordinates from (lam0, theta0, rho0).  nfit > 0 fits on the first nfit
    configurations only; the ratios are evaluated up to the last slice in ts."""
    from scipy.optimize import minimize

    if nfit > 0:
        la = la[:, :nfit]
    C = la.shape[1]
    T = min(T, max(ts) + 1)
    lim = 2.0 ** -6

    def unpack(u):
        lam = float(np.clip(u[0] * 1e-3, -lim, lim))
        theta = float(np.clip(u[1] * 1e-4, -lim, lim))
        rho = float(np.clip(0.96 / (1.0 + np.exp(-u[2])), 0.0, 0.96))
        return lam, theta, rho

    def f(u):
        lr = memory_log_ratio(ctx, seed, config0, C, T, vol, *unpack(u))
        return -memory_log_ess(la, lr, ts)

    r0 = min(max(rho0 / 0.96, 1e-6), 1 - 1e-6)
    u0 = np.array([lam0 * 1e3, theta0 * 1e4, np.log(r0 / (1 - r0))])
    res = minimize(f, u0, method="Nelder-Mead",
                   options={"maxiter": maxiter, "xatol": 1e-2, "fatol": 1e-4,
                            "initial_simplex": [u0, u0 + [0.2, 0, 0], u0 + [0, 0.5, 0],
                                                u0 + [0, 0, 0.3]]})
    return unpack(res.x)


# ------------------------------------------------------------------ the memory network
MEMNET_NP = 35


def memnet_init(lam: float = 0.0, seed: int = 0) -> np.ndarray:
    """Starting parameters of the memory network [W1 (8 x 2), b1 (8), W2 (8), c (2), b2]: the
    constant tilt lam (q(+1) = (1 + tanh lam) / 2, logit 2 lam = b2 2^-10), zero memory
    couplings and small random hidden units (their output weights small, so the network
    starts at the tilt)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MEMNET_NP, np.float32)
    p[:16] = rng.standard_normal(16) * 0.5
    p[16:24] = rng.standard_normal(8) * 0.5
    p[24:32] = rng.standard_normal(8) * 0.05
    p[34] = 2.0 * lam * 1024.0
    return p


def memnet_log_ratio(ctx: Context, seed: int, config0: int, nconf: int, T: int, vol: int,
                     params: np.ndarray) -> np.ndarray:
    """[T][nconf] cumulative log(p / q) of the memory network on the plain sampler's fields
    of configurations config0 .. (nrc_z2_memnet_eval)."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    out = ctx.array((T, nconf), np.float64)
    _check(ctx.lib.nrc_z2_memnet_eval(ctx.h, seed, config0, nconf, T, vol, p.ctypes.data,
                                      out.ptr), "z2_memnet_eval")
    return out.download()


def memnet_objective(la: np.ndarray, lr: np.ndarray, ts):
    """The effective-sample objective of fit_memory (mean over t in ts of
    log E_p[d2]^2 / E_p[d2^2 p / q]) and its weights G[t'][c] = d objective / d sum log q of
    slice t' of configuration c: the mean over t in ts, t > t', of softmax_c(4 x_t + lr_t)."""
    x = la[:, :, 0] if la.ndim == 3 else la
    T = lr.shape[0]
    G = np.zeros((T, x.shape[1]))
    J = 0.0
    for t in ts:
        e = 4 * x[t] + lr[t]
        w = np.exp(e - e.max())
        J += 2 * _lme(2 * x[t]) - _lme(e)
        G[:t] += (w / w.sum())[None, :] / len(ts)
    return J / len(ts), G


def memnet_gradient(ctx: Context, seed: int, config0: int, nconf: int, T: int, vol: int,
                    params: np.ndarray, G: np.ndarray) -> np.ndarray:
    """sum_c grad[c]: the gradient of sum_t G[t][c] sum log q (nrc_z2_memnet_grad)."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    Gd = ctx.to_device(np.ascontiguousarray(G[:T], dtype=np.float64))
    out = ctx.array((nconf, MEMNET_NP), np.float64)
    _check(ctx.lib.nrc_z2_memnet_grad(ctx.h, seed, config0, nconf, T, vol, p.ctypes.data,
                                      Gd.ptr, out.ptr), "z2_memnet_grad")
    return out.download().sum(axis=0)


def fit_memnet(ctx: Context, la: np.ndarray, seed: int, config0: int, T: int, vol: int, ts,
               p0: np.ndarray, iters: int = 120, step: float = 0.05, nfit: int = 0):
    """Trains the memory network on a plain sample (la [T][C] of configurations config0 ..;
    their fields regenerated on the device): Adam ascent of memnet_objective on the first half
    of the configurations (or of the first nfit), the gradient from nrc_z2_memnet_grad; the
    parameters kept are those with the best objective on the other half.  Returns them and a
    record of the training."""
    if nfit > 0:
        la = la[:, :nfit]
    C = la.shape[1]
    T = min(T, max(ts) + 1)
    h = C // 2
    la_tr, la_va = la[:, :h], la[:, h:]
    p = np.asarray(p0, dtype=np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    b1, b2, eps = 0.9, 0.999, 1e-8
    best = (-np.inf, p.astype(np.float32), -1)
    hist = []
    t0 = time.perf_counter()
    for it in range(iters + 1):
        pf = p.astype(np.float32)
        lr_va = memnet_log_ratio(ctx, seed, config0 + h, C - h, T, vol, pf)
        J_va = memnet_objective(la_va, lr_va, ts)[0]
        if J_va > best[0]:
            best = (J_va, pf.copy(), it)
        if it == iters:
            break
        lr_tr = memnet_log_ratio(ctx, seed, config0, h, T, vol, pf)
        J_tr, G = memnet_objective(la_tr, lr_tr, ts)
        g = memnet_gradient(ctx, seed, config0, h, T, vol, pf, G)
        hist.append((it, float(J_tr), float(J_va)))
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mh = m / (1 - b1 ** (it + 1))
        vh = v / (1 - b2 ** (it + 1))
        p = p + step * mh / (np.sqrt(vh) + eps)
    plain = memnet_objective(la_va, np.zeros((T, C - h)), ts)[0]
    rec = {"iterations": iters, "train_configs": h, "validation_configs": C - h,
           "best_iteration": best[2], "validation_objective": float(best[0]),
           "validation_objective_plain": float(plain),
           "seconds": time.perf_counter() - t0,
           "history_every_10": [x for x in hist if x[0] % 10 == 0]}
    return best[1], rec


def _tilt_log_ess(lam: np.ndarray, d2: np.ndarray, S: np.ndarray, V: int):
    """log of the effective sample fraction of the tilted proposal q ~ p prod_t' exp(lam_t'
    S_t') (per-site probability of +1: (1 + tanh lam_t') / 2) for the estimator of <d2>,
    evaluated on a uniform sample: weights w = d2 prod cosh(lam)^V exp(-lam S), so
    E_q[w] = E_p[d2] and E_q[w^2] = E_p[d2^2 prod cosh^V exp(-lam S)].  Returns it and its
    gradient in lam."""
    g = V * np.log(np.cosh(lam))[:, None] - lam[:, None] * S  # [t'][c]
    e = 2.0 * np.log(d2) + g.sum(axis=0)
    m = e.max()
    w = np.exp(e - m)
    log_e2 = np.log(w.mean()) + m
    log_ess = 2.0 * np.log(d2.mean()) - log_e2
    grad = -((V * np.tanh(lam))[:, None] - S) @ (w / w.sum())
    return log_ess, grad


def best_constant_tilt(la: np.ndarray, S: np.ndarray, ts, V: int) -> float:
    """The per-site tilt lambda (q(+1) = (1 + tanh lambda) / 2, logit 2 lambda) that
    maximises the mean over t in ts of the effective sample fraction on the uniform sample
    (la [T][C][m], S [T][C] field sums)."""
    from scipy.optimize import minimize_scalar

    x = la[:, :, 0] if la.ndim == 3 else la

    def f(l):
        tot = 0.0
        for t in ts:
            d2 = np.exp(2.0 * (x[t] - x[t].max()))
            tot += _tilt_log_ess(np.array([l]), d2, S[:t].sum(axis=0, dtype=np.float64)[None],
                                 V * t)[0]
        return -tot / len(ts)

    return float(minimize_scalar(f, bounds=(-0.05, 0.05), method="bounded").x)


def tilt_ceiling(la: np.ndarray, S: np.ndarray, ts, V: int, seed: int = 0) -> dict:
    """The effective sample fraction of the best non-adaptive per-site tilts, measured on a
    uniform sample (la [T][C][m], S [T][C] field sums): for each t in ts, a single tilt for
    every slice before t and one tilt per slice, fitted on one half of the configurations and
    evaluated on the other half (so the fit cannot flatter itself), against the uniform
    sampler's fraction on that half.  The per-sample relative variance is 1 / ess - 1, so
    the variance gain of a tilt at equal cost is (1/ess_uniform - 1) / (1/ess_tilt - 1)."""
    from scipy.optimize import minimize, minimize_scalar

    x = la[:, :, 0] if la.ndim == 3 else la
    C = x.shape[1]
    perm = np.random.default_rng(seed).permutation(C)
    fit, ev = perm[: C // 2], perm[C // 2:]
    out = {}
    for t in ts:
        d2 = np.exp(2.0 * (x[t] - x[t].max()))
        St = S[:t].astype(np.float64)
        Sc = St.sum(axis=0)

        def one(l, idx):
            return _tilt_log_ess(np.array([l]), d2[idx], Sc[None, idx], V * t)[0]

        r1 = minimize_scalar(lambda l: -one(l, fit), bounds=(-0.05, 0.05), method="bounded")
        lam_c = float(r1.x)
        res = minimize(lambda l: tuple(-v for v in _tilt_log_ess(l, d2[fit], St[:, fit], V)),
                       np.full(t, lam_c), jac=True, method="L-BFGS-B")
        ess_u = float(np.exp(_tilt_log_ess(np.zeros(1), d2[ev], Sc[None, ev], V * t)[0]))
        ess_c = float(np.exp(one(lam_c, ev)))
        ess_s = float(np.exp(_tilt_log_ess(res.x, d2[ev], St[:, ev], V)[0]))
        out[int(t)] = {"ess_fraction_uniform": ess_u, "ess_fraction_constant_tilt": ess_c,
                       "ess_fraction_per_slice_tilt": ess_s, "constant_tilt": lam_c,
                       "variance_gain_constant_tilt": (1 / ess_u - 1) / (1 / ess_c - 1),
                       "variance_gain_per_slice_tilt": (1 / ess_u - 1) / (1 / ess_s - 1)}
    return out


def estimator_stats(la: np.ndarray, logr: np.ndarray | None, ts):
    """For each t in ts: log C(t) estimate, relative standard error and the effective sample
    size (sum x)^2 / sum x^2 of x = det(t)^2 r_t (r = 1 for the uniform sampler)."""
    x = la[:, :, 0] if la.ndim == 3 else la
    out = {}
    for t in ts:
        lx = 2.0 * x[t] + (logr[t] if logr is not None else 0.0)
        m = lx.max()
        v = np.exp(lx - m)
        mean = v.mean()
        n = v.size
        rel_se = v.std(ddof=1) / np.sqrt(n) / mean
        out[int(t)] = {"log_C": float(np.log(mean) + m), "rel_se": float(rel_se),
                       "ess": float(v.sum() ** 2 / (v ** 2).sum()), "n": int(n)}
    return out
