#!/usr/bin/env python3
"""Throughput of the auxiliary-field correlator path on one MI355X.

Workload: L=32 lattice (V = 32768), N=33 plane-wave orbitals per spin (the lowest 33
momenta), T=128 slices, tau=0.05, M=1, A=0.4, uniform Z2 fields from Philox, FP32
evolution (FFT + kinetic + field multiplies), FP64 overlaps and LU log-determinants.
A step is one pass of the production slice loop (measure, then evolve; see
production/slater_det/main.C:185-220) over a batch of configurations, i.e. `configs x T`
configuration-slices.  Inputs are generated on the device; the timed region includes
returning the per-configuration log|det| and phase to the host.

Prints one JSON line (rank 0).  `--gpus N` with N > 1 is refused: this build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
LDS_PEAK_GBS = 256 * 128 * 2.4  # LDS: 256 CUs x 128 bytes per clock x 2.4 GHz, GB/s
VALU_SIMDS = 256 * 4  # MI355X: 256 CUs x 4 SIMDs
VALU_CLOCK_HZ = 2.4e9  # peak engine clock


def _tables(L, tau, M):
    """xi_half = 1 + dt/2 * eps(k), eps = (1/2M) sum 4 sin^2(k/2): the reference's Kinetic
    (hamiltonian.C:260-281) with mass M/tau and dt = 1/2, STANDARD dispersion."""
    n = np.arange(L)
    s = np.sin(np.pi * n / L) ** 2
    eps = 2.0 * (s[:, None, None] + s[None, :, None] + s[None, None, :]) / M
    return (1.0 + 0.5 * tau * eps).reshape(-1)


def _harmonic_trap(L, k, dt):
    """q = (V p)^2 for the trapped slice K P I P K, p = exp(-dt v)/V Potential's EXP table
    (hamiltonian.C:163-240): v = (k1 x^2 + k2 y^2 + k3 z^2)/2, coordinates folded to
    [-L/2, L/2)."""
    n = np.arange(L)
    f = np.where(n >= L // 2, n - L, n).astype(float) ** 2
    v = (k[0] * f[:, None, None] + k[1] * f[None, :, None] + k[2] * f[None, None, :]) / 2.0
    return np.exp(-2.0 * dt * v).reshape(-1)


def _lowest(L, N):
    f = np.where(np.arange(L) > L // 2, np.arange(L) - L, np.arange(L))
    pts = [(a, b, c) for a in range(L) for b in range(L) for c in range(L)]
    q = [f[a] ** 2 + f[b] ** 2 + f[c] ** 2 for a, b, c in pts]
    order = sorted(range(len(pts)), key=lambda i: q[i])  # stable: ties keep enumeration order
    return np.array([pts[i] for i in order[:N]], dtype=np.int32)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(L, N, tau, A, configs, slices, threads, reps=3):
    """The FP64 oracle (CPU restatement of the reference algorithm: batched FFTs, N^2
    full-length overlaps, NR LU, every orbital evolved) on `threads` host cores, one
    configuration stream per core as the reference's MPI ranks run (random.C:88): a one-slice
    warm-up per stream, then `reps` barrier-synchronised repetitions of `configs`
    configurations x `slices` slices per stream; the best repetition is the rate."""
    import oracle

    oracle.lib()
    moms = _lowest(L, N)
    kin = _tables(L, tau, 1.0)
    secs = oracle.time_pool(L, N, configs, slices, moms, kin, A, nthreads=threads, reps=reps,
                            config=1 << 20)
    best = float(secs.min())
    cal = fft_calibration(L, N)
    value = threads * configs * slices / best
    # the plane-wave-gather variant (overlaps read at the sink momenta), reported beside the
    # baseline, not as it: one repetition of the same sample
    gsecs = oracle.time_pool(L, N, configs, slices, moms, kin, A, nthreads=threads, reps=1,
                             config=1 << 20, gather=True)
    gather = {"value": threads * configs * slices / float(gsecs.min()),
              "unit": "configs*slices/s",
              "sample": f"the same {threads} streams x {configs} configs x {slices} slices, one "
                        f"repetition ({float(gsecs.min()):.1f} s)",
              "note": "plane-wave sinks measured by reading psi_t at the sink momenta instead of "
                      "N^2 full-length products (the same overlaps); not the reference's path"}
    visible = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = visible
    return {"value": value, "unit": "configs*slices/s", "cores": threads, "kind": "port",
            "sample": f"{threads} streams x {configs} configs x {slices} slices at L={L}, N={N}, "
                      f"FP64, every orbital evolved; 1-slice warm-up, best of {reps} "
                      f"({', '.join(f'{x:.1f}' for x in secs)} s)",
            "cpu_model": _cpu_model(), "cpus_visible": visible, "cpus_in_affinity": affinity,
            "cores_policy": ("one stream per CPU the process may use"
                             if threads >= affinity else
                             f"one stream per core of the host's CPU share for one GPU "
                             f"({threads} of {affinity} CPUs in the affinity mask)"),
            "per_core_value": value / threads,
            "value_if_fft_at_numpy_speed": value * cal["baseline_scale_with_numpy_fft"],
            "value_all_visible_cpus_linear": value / threads * visible,
            "fft_calibration": cal,
            "gather_variant": gather}


def fft_calibration(L, N, reps=3):
    """The oracle's FP64 3D FFT against numpy's (pocketfft) on one core, on the baseline's
    batch (N volumes of L^3, one forward and one backward transform per slice): GFLOP/s at the
    nominal 5 V log2 V per transform, and the baseline's cs/s if its transforms ran at numpy's
    speed (the rest of a slice unchanged)."""
    import oracle

    rng = np.random.default_rng(0)
    a = rng.standard_normal((N, L, L, L)) + 1j * rng.standard_normal((N, L, L, L))
    oracle.fft3d(a, -1)
    np.fft.fftn(a, axes=(1, 2, 3))
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.fft3d(oracle.fft3d(a, -1), 1)
    t_orc = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        np.fft.ifftn(np.fft.fftn(a, axes=(1, 2, 3)), axes=(1, 2, 3))
    t_np = (time.perf_counter() - t0) / reps
    V = L ** 3
    flops = 2 * N * 5.0 * V * np.log2(V)
    moms = _lowest(L, N)
    kin = _tables(L, 0.05, 1.0)
    t_slice = oracle.time_threads(L, N, 2, moms, kin, 0.4, nthreads=1, config=1 << 21) / 2
    return {"oracle_fft_gflops": flops / t_orc / 1e9, "numpy_fft_gflops": flops / t_np / 1e9,
            "ratio": t_np / t_orc, "slice_seconds_1_core": t_slice,
            "fft_seconds_per_slice_1_core": t_orc,
            "baseline_scale_with_numpy_fft": t_slice / max(t_slice - t_orc + t_np, 1e-9),
            "note": "times include the copies of each call; one core"}


def measured_counters(kernel, L, N, C, per_launch, nvo):
    """The committed rocprofv3 counter record of `kernel` on this workload (profiles/
    traffic.json: HBM bytes per launch from FETCH_SIZE / WRITE_SIZE, VALU wave instructions
    per launch), or None if none was taken on this shape.  These are stored figures from the
    profile named in the record, not measured by this run."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            recs = json.load(f)
    except (OSError, ValueError):
        return None
    for t in recs if isinstance(recs, list) else [recs]:
        if (t.get("kernel_id"), t.get("L"), t.get("N"), t.get("configs"),
                t.get("slices_per_launch"), t.get("volumes_per_config")) == (
                kernel, L, N, C, per_launch, nvo):
            return t
    return None


def correlator(la, ph):
    """C(t) = < det(up) det(down) > = < det^2 > over configurations (up and down share the
    field and the orbitals), from log|det| and det/|det|; returned as (log scale, C / scale)."""
    la2 = 2.0 * la[:, :, 0]
    m = la2.max(axis=1, keepdims=True)
    c = (np.exp(la2 - m) * ph[:, :, 0] ** 2).mean(axis=1)
    return m[:, 0], c


ACC_NAMES = {0: "plain FP32", 1: "FP32 state, FP64 x-chain and z-rows",
             2: "FP32 state, FP64 x-chain, FP32 z-rows"}
ACC_KEYS = {0: "plain_fp32", 1: "fp64_x_chain_and_z_rows", 2: "fp64_x_chain"}


def precision_report(ctx, na, L, N, T, tau, A, nconf, seed, conj_modes=(False,), acc=1,
                     ref64=None):
    """Per-t relative error of C(t) from the FP32 path against the FP64 path (every orbital
    evolved) on the same configurations (same Philox fields), and the largest log|det|
    difference; one report per FP32 mode (conjugate partners off / on) in the arithmetic
    `acc` (1: FP64 x-chain and z-rows, 2: FP64 x-chain only, 0: plain FP32).  ref64: the FP64
    run's (logabs, phase), if taken."""
    moms = _lowest(L, N)
    kin = _tables(L, tau, 1.0)
    if ref64 is None:
        e = na.Ensemble(ctx, L, moms, kin, A, nconf, prec=na.PREC_F64, split=na.SPLIT_KSPACE)
        ref64 = e.run(seed, 0, T)[:2]
        e.close()
    la64, ph64 = ref64
    m64, c64 = correlator(la64, ph64)
    reports = []
    for conj in conj_modes:
        e = na.Ensemble(ctx, L, moms, kin, A, nconf, prec=na.PREC_F32, split=na.SPLIT_KSPACE)
        e.set_symmetry(conj)
        e.set_accuracy(acc)
        la32, ph32, _ = e.run(seed, 0, T)
        e.close()
        m32, c32 = correlator(la32, ph32)
        rel = np.abs(c32 * np.exp(m32 - m64) - c64) / np.abs(c64)
        reports.append({
            "configs": nconf, "T": T, "conj_partners": bool(conj),
            "arithmetic": ACC_NAMES[acc],
            "max_rel_err_C": float(rel.max()), "median_rel_err_C": float(np.median(rel)),
            "frac_t_within_1e-6": float((rel <= 1e-6).mean()),
            "rel_err_C_at_t": {str(t): float(rel[t]) for t in (0, T // 4, T // 2, 3 * T // 4,
                                                                T - 1)},
            "max_abs_err_logdet": float(np.abs(la32 - la64).max()),
            "reference": "FP64 evolution of every orbital (FFT, multipliers, overlaps) on the "
                         "same fields"})
    return reports


def batched_uneven(ctx, na, nconf=4096, T=96, tau=0.05, A=0.4, seed=7012, runs=2, conj=True):
    """4096 configurations of uneven shape: L uniform in {16, 24, 32}, N uniform in
    {1, 7, 19, 27, 33} (closed shells), drawn with a fixed seed and grouped by shape; every
    group is autotuned, then all groups run together, T slices captured into one HIP graph.
    Returns aggregate and per-shape configs*slices/s."""
    rng = np.random.default_rng(4096)
    Ls = rng.choice([16, 24, 32], nconf)
    Ns = rng.choice([1, 7, 19, 27, 33], nconf)
    shapes = sorted({(int(l), int(n)) for l, n in zip(Ls, Ns)})
    groups, per_shape = [], []
    for L, N in shapes:
        cnt = int(np.sum((Ls == L) & (Ns == N)))
        e = na.Ensemble(ctx, L, _lowest(L, N), _tables(L, tau, 1.0), A, cnt, prec=na.PREC_F32,
                        split=na.SPLIT_KSPACE)
        e.set_symmetry(conj)
        path, ms = e.tune(3)
        groups.append(e)
        per_shape.append({"L": L, "N": N, "configs": cnt, "volumes_per_config": e.volumes_per_config(),
                          "path": "single-pass" if path == na.PATH_SINGLE_PASS else "two-pass",
                          "tuned_ms_per_slice": ms})
    # each shape alone
    for e, rec in zip(groups, per_shape):
        e.run(seed, 0, T)
        ctx.sync()
        t0 = time.perf_counter()
        e.run(seed, 1 << 20, T)
        rec["cs_per_s_alone"] = rec["configs"] * T / (time.perf_counter() - t0)
    b = na.Batch(ctx, groups)
    out = {}
    for graph in (True, False):
        c0 = [0] * len(groups)
        b.run(seed, c0, T, graph=graph)  # warm-up (and the capture)
        ts = []
        for r in range(runs):
            c0 = [(r + 1) << 20] * len(groups)
            t0 = time.perf_counter()
            b.run(seed, c0, T, graph=graph)
            ts.append(time.perf_counter() - t0)
        out[graph] = (min(ts), b.capture_ms() if graph else 0.0, b.captures())
    b.close()
    for e in groups:
        e.close()
    (tg, cap, ncap), (tn, _, _) = out[True], out[False]
    return {"workload": f"{nconf} configurations, L uniform in {{16,24,32}}, N uniform in "
                        f"{{1,7,19,27,33}}, T={T}, tau={tau}, A={A}, FP32 evolution, FP64 LU",
            "configs": nconf, "T": T, "shape_groups": len(groups), "conj_partners": bool(conj),
            "value": nconf * T / tg, "unit": "configs*slices/s",
            "ms_per_run": 1e3 * tg, "graph": True,
            "capture_and_instantiate_ms": cap,
            "value_without_graph": nconf * T / tn,
            "graph_captures": ncap,
            "timing": "one graph launch per run; the graph is captured once, in the warm-up "
                      "run, and the timed runs (other configuration ranges) reuse it",
            "per_shape": per_shape}


def learned_proposal(ctx, na, L=24, N=14, T=96, tau=0.05, A=0.4, n_train=16384, n_plain=65536,
                     n_learned=16384, steps=300, batch=64, lr=4e-3, seed=7012, optimizer=0,
                     out_scale=0.0, conj=True, n_fit=4096, n_memory=65536, mn_iters=120):
    """Effective samples per second of the learned proposal against the uniform sampler.
    The uniform sampler generates n_train configurations; the network is trained on them
    (weighted maximum likelihood, det^2 weights); then n_plain uniform and n_learned proposed
    configurations estimate C(t) = <det(t)^2> (n_memory with fields with memory, whose three
    parameters are fitted on the first n_fit training configurations).  Per t in
    {T/4, T/2, 3T/4}: the relative standard error of each estimate, the per-sample variance
    v = n rel_se^2 and the cost per sample tau give the effective samples per second
    1 / (v tau) at equal error."""
    from nrcps_amd import proposal as pr

    moms, kin = _lowest(L, N), _tables(L, tau, 1.0)
    ts = [T // 4, T // 2, 3 * T // 4]
    chunk = min(n_train, n_learned, 16384)
    ens = na.Ensemble(ctx, L, moms, kin, A, chunk, prec=na.PREC_F32, split=na.SPLIT_KSPACE)
    ens.set_symmetry(conj)  # the same evolution engine for both samplers
    ens.run(seed, 0, 2)  # warm-up
    # uniform sampler: the training sample, then the comparison sample
    t0 = time.perf_counter()
    la_train = np.concatenate([ens.run(seed, k, T)[0] for k in range(0, n_train, chunk)], axis=1)
    t_train_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    la_plain = np.concatenate([ens.run(seed, (1 << 22) + k, T)[0]
                               for k in range(0, n_plain, chunk)], axis=1)
    t_plain = time.perf_counter() - t0
    # training: the output bias starts at the best constant tilt of the training sample (one
    # parameter fitted to its fields and log-determinants), then the network trains
    t0 = time.perf_counter()
    V = L ** 3
    S_train = np.concatenate([pr.z2_sums(ctx, seed, k, chunk, T, V)
                              for k in range(0, n_train, chunk)], axis=1)
    params = pr.init_params(np.random.default_rng(0), out_scale=out_scale)
    lam_fit = pr.best_constant_tilt(la_train, S_train, ts, V)
    params[8240] = 2.0 * lam_fit
    prop = pr.Proposal(ctx, L, max(batch, chunk), params)
    losses, mu, sd = pr.train(ctx, prop, la_train, seed, 0, steps, batch, lr, ts,
                              np.random.default_rng(1), optimizer=optimizer)
    ctx.sync()
    t_train = time.perf_counter() - t0
    # the measured ceiling: the best non-adaptive per-site tilts on the uniform comparison
    # sample (fitted on one half, evaluated on the other)
    S_plain = np.concatenate([pr.z2_sums(ctx, seed, (1 << 22) + k, chunk, T, V)
                              for k in range(0, n_plain, chunk)], axis=1)
    ceiling = pr.tilt_ceiling(la_plain, S_plain, ts, V)
    # proposed configurations
    sampler = pr.LearnedSampler(ctx, prop, ens, mu, sd)
    sampler.run(seed + 1, 1 << 23, 2)  # warm-up
    ctx.sync()
    t0 = time.perf_counter()
    parts = [sampler.run(seed + 1, (1 << 24) + k, T) for k in range(0, n_learned, chunk)]
    t_learned = time.perf_counter() - t0
    la_l = np.concatenate([p[0] for p in parts], axis=1)
    logr = np.concatenate([p[2] for p in parts], axis=1)
    # the same network drawing four slices' fields per group, evolved in four-slice launches
    sampler.run(seed + 3, 1 << 23, 4, group=4)  # warm-up
    ctx.sync()
    t0 = time.perf_counter()
    gparts = [sampler.run(seed + 3, (1 << 26) + k, T, group=4) for k in range(0, n_learned, chunk)]
    t_grouped = time.perf_counter() - t0
    la_g = np.concatenate([q[0] for q in gparts], axis=1)
    logr_g = np.concatenate([q[2] for q in gparts], axis=1)
    # the fitted constant tilt alone: tilted Philox fields on the uniform sampler's engine
    q = ens.set_tilt(0.5 * (1.0 + np.tanh(lam_fit)))
    ens.run(seed + 2, 1 << 23, 2)  # warm-up
    t0 = time.perf_counter()
    tparts = []
    for k in range(0, n_learned, chunk):
        la_k = ens.run(seed + 2, (1 << 25) + k, T)[0]
        tparts.append((la_k, ens.log_ratio(T)))
    t_tilt = time.perf_counter() - t0
    ens.set_tilt(0.5)
    la_t = np.concatenate([a for a, _ in tparts], axis=1)
    logr_t = np.concatenate([b for _, b in tparts], axis=1)
    # fields with memory: (lambda, theta, rho) fitted to the training sample (its fields
    # regenerated on the device, the proposal's exact ratio evaluated on them), then drawn on
    # the uniform sampler's engine and reweighted exactly
    t0 = time.perf_counter()
    mem = pr.fit_memory(ctx, la_train, seed, 0, T, V, ts, lam0=float(np.tanh(lam_fit)),
                        nfit=n_fit)
    t_mem_fit = time.perf_counter() - t0
    mem = ens.set_memory(*mem)
    ens.run(seed + 4, 1 << 23, 2)  # warm-up
    ctx.sync()
    t0 = time.perf_counter()
    mparts = []
    for k in range(0, n_memory, chunk):
        la_k = ens.run(seed + 4, (1 << 27) + k, T)[0]
        mparts.append((la_k, ens.log_ratio(T)))
    t_mem = time.perf_counter() - t0
    ens.set_memory(0.0, 0.0, 0.5)
    la_m = np.concatenate([a for a, _ in mparts], axis=1)
    logr_m = np.concatenate([b for _, b in mparts], axis=1)
    # the memory network: trained on the training sample (its fields regenerated on the
    # device; Adam on the exact effective-sample objective, the parameters kept by the best
    # objective on a held-out half), drawn on the uniform sampler's engine, reweighted exactly
    t0 = time.perf_counter()
    mn_p, mn_rec = pr.fit_memnet(ctx, la_train, seed, 0, T, V, ts,
                                 pr.memnet_init(float(np.tanh(lam_fit)), 0), iters=mn_iters,
                                 nfit=n_fit)
    t_mn_fit = time.perf_counter() - t0
    ens.set_memnet(mn_p)
    ens.run(seed + 5, 1 << 23, 2)  # warm-up
    ctx.sync()
    t0 = time.perf_counter()
    nparts = []
    for k in range(0, n_memory, chunk):
        la_k = ens.run(seed + 5, (1 << 28) + k, T)[0]
        nparts.append((la_k, ens.log_ratio(T)))
    t_mn = time.perf_counter() - t0
    ens.set_memnet(None)
    la_n = np.concatenate([a for a, _ in nparts], axis=1)
    logr_n = np.concatenate([b for _, b in nparts], axis=1)
    sn = pr.estimator_stats(la_n, logr_n, ts)
    sm = pr.estimator_stats(la_m, logr_m, ts)
    sp = pr.estimator_stats(la_plain, None, ts)
    sl = pr.estimator_stats(la_l, logr, ts)
    st = pr.estimator_stats(la_t, logr_t, ts)
    sg = pr.estimator_stats(la_g, logr_g, ts)
    tau_p, tau_l = t_plain / n_plain, t_learned / n_learned
    tau_t = t_tilt / n_learned
    tau_g = t_grouped / n_learned
    tau_m = t_mem / n_memory
    tau_m_amort = (t_mem + t_mem_fit + t_train_gen) / n_memory
    # the uniform training sample is itself a plain sample of C(t): reused, the run's estimate
    # is the inverse-variance combination of it and the memory sample, bought with the
    # training sample, the fit and the memory sample together
    st_train = pr.estimator_stats(la_train, None, ts)
    t_all = t_train_gen + t_mem_fit + t_mem
    memory = {}
    for t in ts:
        vp = sp[t]["rel_se"] ** 2 * n_plain
        vm = sm[t]["rel_se"] ** 2 * n_memory
        vtr = st_train[t]["rel_se"] ** 2 * n_train
        sig = float(np.hypot(sp[t]["rel_se"], sm[t]["rel_se"]))
        info = n_train / vtr + n_memory / vm  # inverse variance of the combined estimate
        memory[str(t)] = {"log_C": sm[t]["log_C"], "rel_se": sm[t]["rel_se"],
                          "bias_within_2sigma": bool(abs(sm[t]["log_C"] - sp[t]["log_C"])
                                                     <= 2 * sig),
                          "variance_gain": vp / vm, "gain": (vp * tau_p) / (vm * tau_m),
                          "gain_with_fit_amortised": (vp * tau_p) / (vm * tau_m_amort),
                          "gain_with_fit_amortised_training_sample_reused":
                              info * vp * tau_p / t_all}
    # configurations of the memory sampler after which its gain with the one-time cost (the
    # uniform training sample and the fit) amortised reaches 3x: g n tau / (n tau + F) = 3
    g_min = min(v["gain"] for v in memory.values())
    mem_3x = (3.0 * (t_mem_fit + t_train_gen) / (tau_m * (g_min - 3.0))
              if g_min > 3.0 else None)
    # the same with the training sample reused: (n_train + n g_v) tau_p = 3 (F + n tau_m) at
    # the smallest variance gain g_v over t
    gv_min = min(v["variance_gain"] for v in memory.values())
    den = gv_min * tau_p - 3.0 * tau_m
    mem_3x_reused = (max(0.0, (3.0 * (t_train_gen + t_mem_fit) - n_train * tau_p) / den)
                     if den > 0 else None)
    tau_n = t_mn / n_memory
    tau_n_amort = (t_mn + t_mn_fit + t_train_gen) / n_memory
    t_all_n = t_train_gen + t_mn_fit + t_mn
    memnet = {}
    for t in ts:
        vp = sp[t]["rel_se"] ** 2 * n_plain
        vn = sn[t]["rel_se"] ** 2 * n_memory
        vtr = st_train[t]["rel_se"] ** 2 * n_train
        sig = float(np.hypot(sp[t]["rel_se"], sn[t]["rel_se"]))
        info = n_train / vtr + n_memory / vn
        memnet[str(t)] = {"log_C": sn[t]["log_C"], "rel_se": sn[t]["rel_se"],
                          "bias_within_2sigma": bool(abs(sn[t]["log_C"] - sp[t]["log_C"])
                                                     <= 2 * sig),
                          "variance_gain": vp / vn, "gain": (vp * tau_p) / (vn * tau_n),
                          "gain_with_fit_amortised": (vp * tau_p) / (vn * tau_n_amort),
                          "gain_with_fit_amortised_training_sample_reused":
                              info * vp * tau_p / t_all_n}
    grouped = {}
    for t in ts:
        vp = sp[t]["rel_se"] ** 2 * n_plain
        vg = sg[t]["rel_se"] ** 2 * n_learned
        sig = float(np.hypot(sp[t]["rel_se"], sg[t]["rel_se"]))
        grouped[str(t)] = {"log_C": sg[t]["log_C"], "rel_se": sg[t]["rel_se"],
                           "bias_within_2sigma": bool(abs(sg[t]["log_C"] - sp[t]["log_C"])
                                                      <= 2 * sig),
                           "variance_gain": vp / vg, "gain": (vp * tau_p) / (vg * tau_g)}
    tilted = {}
    for t in ts:
        vp = sp[t]["rel_se"] ** 2 * n_plain
        vt = st[t]["rel_se"] ** 2 * n_learned
        sig = float(np.hypot(sp[t]["rel_se"], st[t]["rel_se"]))
        tilted[str(t)] = {"log_C": st[t]["log_C"], "rel_se": st[t]["rel_se"],
                          "bias_within_2sigma": bool(abs(st[t]["log_C"] - sp[t]["log_C"])
                                                     <= 2 * sig),
                          "variance_gain": vp / vt, "gain": (vp * tau_p) / (vt * tau_t)}
    tau_l_amort = (t_learned + t_train + t_train_gen) / n_learned
    per_t = {}
    for t in ts:
        vp = sp[t]["rel_se"] ** 2 * n_plain
        vl = sl[t]["rel_se"] ** 2 * n_learned
        sig = float(np.hypot(sp[t]["rel_se"], sl[t]["rel_se"]))
        per_t[str(t)] = {
            "log_C_plain": sp[t]["log_C"], "log_C_learned": sl[t]["log_C"],
            "rel_se_plain": sp[t]["rel_se"], "rel_se_learned": sl[t]["rel_se"],
            "bias_within_2sigma": bool(abs(sl[t]["log_C"] - sp[t]["log_C"]) <= 2 * sig),
            "ess_plain": sp[t]["ess"], "ess_learned": sl[t]["ess"],
            "eff_samples_per_s_plain": 1.0 / (vp * tau_p),
            "eff_samples_per_s_learned": 1.0 / (vl * tau_l),
            "gain": (vp * tau_p) / (vl * tau_l),
            "gain_with_training_amortised": (vp * tau_p) / (vl * tau_l_amort),
            "variance_gain": vp / vl,
            "ceiling": ceiling[t]}
    ps = prop.get_params()
    prop.close()
    ens_nvo = ens.volumes_per_config()
    ens.close()
    w = np.exp(logr[ts[-1]] - logr[ts[-1]].max())  # the ratios r = prod 0.5 / q alone
    return {"workload": f"L={L}, N={N}, T={T}, tau={tau}, A={A}; {n_train} uniform "
                        f"configurations for training, {n_plain} uniform, {n_memory} with "
                        f"memory and {n_learned} from each other proposal for C(t)",
            "metric": "effective samples/s at equal statistical error of C(t)",
            "conj_partners": bool(conj), "volumes_per_config": ens_nvo,
            "best_learned_proposal": max(
                ({"proposal": name,
                  "gain_min_over_t": min(v["gain"] for v in rec.values()),
                  "gain_with_fit_amortised_training_sample_reused_min_over_t": min(
                      v["gain_with_fit_amortised_training_sample_reused"]
                      for v in rec.values()),
                  "bias_within_2sigma_every_t": all(v["bias_within_2sigma"]
                                                    for v in rec.values())}
                 for name, rec in (("memory_fields", memory), ("memory_network", memnet))),
                key=lambda r: r["gain_min_over_t"]),
            "memory_network": {
                "what": "a pointwise network (8 ReLU units and a linear skip, 35 parameters) "
                        "over each site's two decayed memories of its own past field (decays "
                        "0.75, 0.94; nrc_ensemble_set_memnet), trained by Adam on the exact "
                        "effective-sample objective of the training sample with the project's "
                        "own evaluation and gradient kernels (nrc_z2_memnet_eval, "
                        "nrc_z2_memnet_grad), drawn on the uniform sampler's kernels and "
                        "reweighted exactly",
                "cost_ratio_to_uniform": tau_n / tau_p,
                "gain_min_over_t": min(v["gain"] for v in memnet.values()),
                "gain_with_fit_amortised_min_over_t": min(v["gain_with_fit_amortised"]
                                                          for v in memnet.values()),
                "gain_with_fit_amortised_training_sample_reused_min_over_t": min(
                    v["gain_with_fit_amortised_training_sample_reused"]
                    for v in memnet.values()),
                "training": mn_rec, "params": [float(x) for x in mn_p],
                "configs": n_memory, "per_t": memnet, "seconds": t_mn,
                "training_seconds": t_mn_fit},
            "network_gain_min_over_t": min(v["gain"] for v in per_t.values()),
            "network_cost_ratio_to_uniform": tau_l / tau_p,
            "ceiling_note": "variance gain of the best non-adaptive per-site tilt (constant, or "
                            "one per slice) measured on the uniform sample: the gain in "
                            "effective samples/s such a proposal reaches at the cost of a "
                            "uniform configuration; the learned proposal's gain is its "
                            "variance gain over its cost ratio",
            "network_per_t": per_t,
            "network_grouped": {
                "what": "the trained network drawing four slices' fields in a row (each on the "
                        "previous field and the log-weight of the last row measured before "
                        "the group), evolved in four-slice launches (nrc_ensemble_evolve_fused)",
                "cost_ratio_to_uniform": tau_g / tau_p,
                "gain_min_over_t": min(v["gain"] for v in grouped.values()),
                "per_t": grouped, "seconds": t_grouped},
            "memory_fields": {
                "what": "fields with memory: each site +1 with probability (1 + lambda + theta m)"
                        " / 2, m = sum_k rho^(k-1) phi(t-k) its own decayed past field "
                        "(nrc_ensemble_set_memory), fitted to the training sample through the "
                        "exact ratio evaluated on its fields (nrc_z2_memory_eval), drawn on the "
                        "uniform sampler's kernels and reweighted exactly",
                "lambda": mem[0], "theta": mem[1], "rho": mem[2],
                "cost_ratio_to_uniform": tau_m / tau_p,
                "gain_min_over_t": min(v["gain"] for v in memory.values()),
                "gain_with_fit_amortised_min_over_t": min(v["gain_with_fit_amortised"]
                                                          for v in memory.values()),
                "gain_with_fit_amortised_training_sample_reused_min_over_t": min(
                    v["gain_with_fit_amortised_training_sample_reused"]
                    for v in memory.values()),
                "configs_for_3x_with_fit_amortised": mem_3x,
                "configs_for_3x_with_fit_amortised_training_sample_reused": mem_3x_reused,
                "fit_configs": min(n_fit, n_train) if n_fit > 0 else n_train,
                "configs": n_memory,
                "per_t": memory, "seconds": t_mem, "fit_seconds": t_mem_fit},
            "fitted_tilt": {
                "what": "the constant per-site tilt fitted to the training sample's field sums "
                        "and log-determinants, drawn as tilted Philox fields on the uniform "
                        "sampler's kernels (nrc_ensemble_set_tilt) and reweighted exactly",
                "lambda": lam_fit, "p_plus": q, "cost_ratio_to_uniform": tau_t / tau_p,
                "gain_min_over_t": min(v["gain"] for v in tilted.values()),
                "per_t": tilted, "seconds": t_tilt},
            "proposal_weight_ess_fraction_at_3T/4": float(w.sum() ** 2 / (w ** 2).sum() / w.size),
            "seconds": {"uniform_training_sample": t_train_gen, "training": t_train,
                        "uniform_comparison_sample": t_plain, "proposed_sample": t_learned,
                        "tilted_sample": t_tilt, "grouped_proposed_sample": t_grouped,
                        "memory_fit": t_mem_fit, "memory_sample": t_mem,
                        "memory_network_training": t_mn_fit,
                        "memory_network_sample": t_mn},
            "training": {"steps": steps, "batch": batch, "lr": lr,
                         "optimizer": "sgd-momentum" if optimizer == 1 else "adam",
                         "loss_first": float(losses[0]), "loss_last10": float(losses[-10:].mean()),
                         "output_bias": float(ps[8240]),
                         # the output bias starts at the fitted constant tilt, so the first
                         # loss already holds what a constant output gains over q = 1/2:
                         # KL(1/2 + d || 1/2) ~ 2 d^2 nats per site, d = tanh(lambda) / 2, a
                         # drop in the seventh decimal for a 2x variance gain; training adds
                         # only what the fields' spatial structure and u carry beyond it
                         "loss_gain_of_the_fitted_tilt_over_q_half_per_site":
                             2.0 * (0.5 * float(np.tanh(lam_fit))) ** 2,
                         "loss_drop_in_training": float(losses[0] - losses[-10:].mean())}}


def dense_overlap(ctx, na, L=32, N=33, T=128, C=1024, tau=0.05, A=0.4, seed=7012, ref64=None,
                  acc=1):
    """The overlap matrix Phi^H Psi(t) as a batched complex GEMM on the matrix cores (dense
    sinks, OneBody::Project / Cuda::InnerProduct for any sink vectors): split-BF16 products
    with FP32 accumulation; with the default arithmetic (acc 1) the splits are three-way
    (BF16x6: six products, FP32-exact inputs) and the evolution runs the FP64 x-chain four
    slices per launch, writing the folded states inside a launch for the GEMM (which reads the
    launch's first state with its x-DFT pending, against sinks transformed to match); acc 0:
    split-BF16x3 and plain FP32 single-slice evolution.  With
    conjugate partners the stored volumes (one per (k, -k) pair) are projected on the sinks
    and on their conjugate mirrors, two GEMMs of N sink rows each; every orbital evolved (the
    reference's layout) is reported beside it.  The sinks are the plane waves as dense k-space
    vectors, so the gathered measurement is the check.  Reports the overlap kernels' time per
    slice against the MFMA peak (the executed terms x 8 N n_cols V bf16 flops per
    configuration-slice, n_cols = 2 n_vol with partners, N without) and the HBM peak (8 n_cols
    V bytes read per configuration-slice)."""
    moms, kin = _lowest(L, N), _tables(L, tau, 1.0)
    V = L ** 3
    sinks = np.zeros((N, V), np.complex128)
    sinks[np.arange(N), (moms[:, 0] * L + moms[:, 1]) * L + moms[:, 2]] = 1.0
    terms = 6 if acc != 0 else 3

    def run(conj):
        e = na.Ensemble(ctx, L, moms, kin, A, C, prec=na.PREC_F32, split=na.SPLIT_KSPACE,
                        sinks=sinks)
        e.set_path(na.PATH_SINGLE_PASS)
        e.set_accuracy(acc)
        e.set_symmetry(conj)
        nvo = e.volumes_per_config()
        e.tune(8)  # GEMM variants for the shapes, slices per launch
        tuned = [(k, v, round(ms, 4)) for k, v, ms in e.tune_report()]
        e.run(seed, 0, 2)
        ctx.sync()
        t0 = time.perf_counter()
        la, ph, _ = e.run(seed, 0, T)
        ctx.sync()
        dt = time.perf_counter() - t0
        e.profile(True)
        e.run(seed, 0, T)
        t_ovl, n_ovl = e.kernel_time(4)
        e.profile(False)
        e.close()
        ms = t_ovl / max(n_ovl, 1)
        cols = 2 * nvo if conj else N
        flops = terms * 8.0 * N * cols * V * C
        byts = 8.0 * cols * V * C
        return la, ph, {
            "conj_partners": conj, "volumes_per_config": nvo, "tuner": tuned,
            "value": C * T / dt, "unit": "configs*slices/s", "overlap_ms_per_slice": ms,
            "mfma": {"achieved_tflops": flops / (ms * 1e-3) / 1e12, "peak_tflops": 2500.0,
                     "frac": flops / (ms * 1e-3) / 2.5e15},
            "hbm": {"achieved_gbs": byts / (ms * 1e-3) / 1e9, "peak_gbs": HBM_PEAK_GBS,
                    "frac": byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}

    la_d, ph_d, rec = run(True)
    la_f, ph_f, rec_full = run(False)
    g = na.Ensemble(ctx, L, moms, kin, A, C, prec=na.PREC_F32, split=na.SPLIT_KSPACE)
    la_g, _, _ = g.run(seed, 0, T)
    g.close()

    def prec_of(la, ph, what):
        if ref64 is None or ref64[0].shape[:2] != (T, C):
            return None
        m64, c64 = correlator(*ref64)
        m32, c32 = correlator(la, ph)
        rel = np.abs(c32 * np.exp(m32 - m64) - c64) / np.abs(c64)
        return {"max_rel_err_C": float(rel.max()), "median_rel_err_C": float(np.median(rel)),
                "frac_t_within_1e-6": float((rel <= 1e-6).mean()),
                "reference": "FP64 path on the same fields (the precision report's)",
                "arithmetic": (f"FP32 evolution with the FP64 x-chain (four slices per launch, "
                               f"the states inside a launch written for the GEMM, {what}), "
                               f"split-BF16x6 overlaps on MFMA with FP32 accumulation, FP64 LU"
                               if acc != 0 else
                               f"plain FP32 evolution (single-slice kernel, {what}), "
                               f"split-BF16x3 overlaps on MFMA with FP32 accumulation, FP64 LU")}

    rec["precision"] = prec_of(la_d, ph_d, "conjugate partners")
    rec_full["precision"] = prec_of(la_f, ph_f, "every orbital")
    rec["max_abs_diff_logdet_vs_gather"] = float(np.abs(la_d - la_g).max())
    rec_full["max_abs_diff_logdet_vs_gather"] = float(np.abs(la_f - la_g).max())
    return {"workload": f"dense sinks (the precision form as written): L={L}, N={N}, T={T}, "
                        f"{C} configurations, FP32 evolution, split-BF16x{terms} overlaps on "
                        f"MFMA, FP64 LU",
            "terms": terms, **rec, "every_orbital": rec_full,
            "timing": "HIP events around the overlap kernels of each measurement"}


def fp64_config2(ctx, na, L=16, N=7, T=64, C=2048, tau=0.05, A=0.4, seed=7012, threads=0):
    """The plain FP64 GPU path at the FP64 parity workload (L=16, N=7 plane waves, T=64, 2048
    configurations, Philox Z2 fields): configs*slices/s with every orbital evolved (the
    reference's layout) and with conjugate partners, the largest |d log|det|| of the first
    configurations against the FP64 oracle on the same fields, and (threads > 0) the oracle's
    own rate at this shape on that many host cores, one configuration stream per core."""
    import oracle

    moms, kin = _lowest(L, N), _tables(L, tau, 1.0)
    rec = {"workload": f"L={L}, N={N}, T={T}, tau={tau}, A={A}, {C} configurations, FP64 "
                       f"evolution, plane-wave sinks gathered, FP64 LU",
           "unit": "configs*slices/s"}
    la0 = None
    for conj in (False, True):
        e = na.Ensemble(ctx, L, moms, kin, A, C, prec=na.PREC_F64, split=na.SPLIT_KSPACE)
        e.set_symmetry(conj)
        e.run(seed, 1 << 20, 2)
        ctx.sync()
        t0 = time.perf_counter()
        la, _, _ = e.run(seed, 0, T)
        ctx.sync()
        dt = time.perf_counter() - t0
        e.close()
        rec["value_conj_partners" if conj else "value"] = C * T / dt
        if not conj:
            la0 = la
    oracle.lib()
    err = 0.0
    for c in range(2):
        ref = oracle.run_config(L, N, T, moms, kin, A=A, seed=seed, config=c)
        err = max(err, float(np.abs(la0[:, c, 0] - ref["logabs"][:, 0]).max()))
    rec["max_abs_diff_logdet_vs_oracle_first_2_configs"] = err
    if threads > 0:
        secs = oracle.time_pool(L, N, 2, T, moms, kin, A, nthreads=threads, reps=1,
                                config=1 << 20)
        rec["cpu_oracle"] = {"value": threads * 2 * T / float(secs.min()), "cores": threads,
                             "sample": f"{threads} streams x 2 configs x {T} slices, FP64, "
                                       f"every orbital"}
        rec["speedup_vs_cpu_oracle"] = rec["value"] / rec["cpu_oracle"]["value"]
    return rec


def pair_sinks(ctx, na, L=32, N=35, T=32, C=256, tau=0.05, A=0.4, seed=7012, conj=True,
               trap=True):
    """The pair-sink measurement of production/slater_det3_trapped (L=32, N=35): every slice
    the pair matrix phi[s][t] = sum_p w(p) psi_s(p) psi_t(-p) of every configuration (split-
    BF16x3 on MFMA) and SlaterDet3's 2N determinants, plane-wave sources, the PAIR1 wave
    function w = 1/(p^2 + 0.5).  Reports configs*slices/s and the pair-matrix kernel against
    the MFMA peak (3 x 8 N^2 V flops per configuration-slice) and its algorithmic HBM bytes
    (the orbitals read once: 8 N V)."""
    n = np.arange(L)
    f = np.where(n > L // 2, n - L, n) * (2 * np.pi / L)
    p2 = (f[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2).reshape(-1)
    w = 1.0 / (p2 + 0.5)
    m = (-n) % L
    opp = ((m[:, None, None] * L + m[None, :, None]) * L + m[None, None, :]).reshape(-1)
    ens = na.Ensemble(ctx, L, _lowest(L, N), _tables(L, tau, 1.0), A, C, prec=na.PREC_F32,
                      split=na.SPLIT_KSPACE, pair_w=w, pair_opp=opp.astype(np.int32))
    ens.set_symmetry(conj)
    nvo = ens.volumes_per_config()
    trapped = None
    if trap:  # the production trap: K P I P K with potential.arg's harmonic EXP potential
        ens.set_trap(_harmonic_trap(L, (0.000625, 0.000625, 0.000625), 0.5))
        ens.run(seed, 0, 2)
        ctx.sync()
        t0 = time.perf_counter()
        la_t, _, _ = ens.run(seed, C, T)
        ctx.sync()
        dtt = time.perf_counter() - t0
        trapped = {"value": C * T / dtt, "unit": "configs*slices/s",
                   "finite_fraction": float(np.isfinite(la_t).mean()),
                   "trap": "harmonic, EXP, spring constants 0.000625 (slater_det3_trapped's "
                           "potential.arg), K P I P K per slice, the trap applied inside the "
                           "FP64 x-chain kernel"}
        ens.set_trap(None)
    ens.run(seed, 0, 2)
    ctx.sync()
    t0 = time.perf_counter()
    la, _, _ = ens.run(seed, C, T)
    ctx.sync()
    dt = time.perf_counter() - t0
    ens.profile(True)
    ens.run(seed, 2 * C, T)
    t_evo, n_evo = ens.kernel_time(0)
    t_lu, n_lu = ens.kernel_time(2)
    t_meas, n_meas = ens.kernel_time(4)
    ens.profile(False)
    ens.close()
    V = L ** 3
    meas_ms = t_meas / max(n_meas, 1)  # gather + pair matrices of one slice
    # products computed: N x N pair matrices, or with conjugate partners two nvo x nvo ones
    # (mirrored sites, and the same sites conjugated) over the nvo stored volumes
    prods = 2 * nvo * nvo if nvo < N else N * N
    flops = 3 * 8.0 * prods * V * C
    return {"workload": f"pair sinks (SlaterDet3): L={L}, N={N}, T={T}, {C} configurations, "
                        f"FP32 evolution, split-BF16x3 pair matrices, 2N FP64 LU per slice",
            "value": C * T / dt, "unit": "configs*slices/s",
            "conj_partners": bool(conj), "volumes_per_config": nvo, "trapped": trapped,
            "finite_fraction": float(np.isfinite(la).mean()),
            "finite_fraction_t_ge_1": float(np.isfinite(la[1:]).mean()),
            "finite_note": "log|det| = -inf marks a pair matrix with an all-zero row, not a "
                           "numerical failure: at t = 0 the sources are plane waves, and the "
                           "pair overlaps sum_p w(p) psi_a(p) psi_b(-p) vanish unless k_a = "
                           "-k_b, so minors without complete (k, -k) pairs have zero rows "
                           "(src/utils/ludcmp.C:25-28 stops the program on such a row; the "
                           "batched LU reports it as a status and -inf instead); from t = 1 "
                           "on every value is finite",
            "avg_ms_per_slice": {"evolution": t_evo / T,
                                 "pair_matrices_and_gather": meas_ms,
                                 "lu_2N": t_lu / max(n_meas, 1)},
            "evolution_launches_per_run": n_evo,
            "pair_matrix_mfma": {"achieved_tflops": flops / (meas_ms * 1e-3) / 1e12,
                                 "peak_tflops": 2500.0,
                                 "frac": flops / (meas_ms * 1e-3) / 1e12 / 2500.0},
            "pair_matrix_hbm": {"algorithmic_gbs": 8.0 * nvo * V * C / (meas_ms * 1e-3) / 1e9,
                                "peak_gbs": HBM_PEAK_GBS,
                                "frac": 8.0 * nvo * V * C / (meas_ms * 1e-3) / 1e9
                                        / HBM_PEAK_GBS,
                                "bound": "hbm: the stored orbitals are read once per slice, "
                                         "114 useful bf16 flops per byte"}}


def evolution_roofline(L, N, T, C, V, nvo, args, tx, nx, tyz, nyz, tlu, nlu, tf, nf, tg, ng):
    """The roofline record of the evolution kernel of the headline run.  Algorithmic bytes per
    configuration-slice: read + write the evolved FP32 volumes once plus phi as one byte per
    site (B = nvo 16 V + V); achieved = B x configs x slices per launch / launch time (HIP
    events on the ensemble's stream).  The kernel is bound by its VALU issue: the committed
    counter record of the same kernel and shape (SQ_INSTS_VALU per launch) over this run's
    launch time gives the issue rate against 614 G wave instructions/s (256 CUs x 4 SIMDs x
    2.4 GHz / 4 cycles per wave64 instruction, FP32 and FP64 alike), and its FETCH_SIZE /
    WRITE_SIZE the HBM bytes actually moved."""
    b_evo = nvo * 16 * V + V
    single = nyz == 0
    per_launch = T / max(nx, 1)  # slices per launch of the evolution kernel
    launch_ms = tx / max(nx, 1)
    if single and args.acc != 0 and args.path != 1:
        kid, kernel = "k_slicex", f"k_slicex ({per_launch:g} slices per launch, FP64 x-chain)"
    elif single and per_launch > 1:
        kid, kernel = "k_slice1p", f"k_slice1p ({per_launch:g} slices per launch)"
    elif single:
        kid, kernel = "k_slice1", "k_slice1 (one launch per slice)"
    else:
        kid, kernel = "k_xpass+k_yzpass", "k_xpass + k_yzpass (one slice)"
    alg_bytes = b_evo * C * per_launch
    slice_ms = (tx + tyz) / T
    hbm_alg = b_evo * C / (slice_ms * 1e-3) / 1e9
    cnt = measured_counters(kid, L, N, C, per_launch, nvo) if single else None
    valu_peak = VALU_SIMDS * VALU_CLOCK_HZ / 4.0
    rec = {"kernel": kernel,
           "hbm_algorithmic": {"achieved_gbs": hbm_alg, "peak_gbs": HBM_PEAK_GBS,
                               "frac": hbm_alg / HBM_PEAK_GBS},
           "traffic": cnt["traffic_bytes_per_launch"] if cnt else None,
           "traffic_source": (cnt["source"] + " (stored counter record, not measured by "
                              "this run)") if cnt else None,
           "algorithmic_bytes_per_launch": alg_bytes,
           "avg_ms": {"evolution_per_slice": slice_ms, "evolution_launch": launch_ms,
                      "yzpass": tyz / max(nyz, 1), "lu": tlu / max(nlu, 1),
                      "fields": tf / max(nf, 1), "gather": tg / max(ng, 1)},
           "bytes_per_cs": b_evo, "volumes_per_config": nvo,
           "timing": "HIP events on the ensemble's stream around every evolution launch"}
    if kid == "k_slicex" and L == 32:
        # LDS bytes per volume-slice of k_slicex at L = 32 (two x-pencils per thread, two
        # rounds of 16 planes): the parked pencil written, swapped and restored (4 x V/2 x 8),
        # per round the planes written and read back and the z rows, y columns and z rows read
        # and written (8 x V/2 x 8 per round, two rounds), the field word of every site (4 V)
        lds = (16 + 64 + 4) * V * C * nvo * per_launch
        lds_gbs = lds / (launch_ms * 1e-3) / 1e9
        rec["lds_algorithmic"] = {"bytes_per_launch": lds, "achieved_gbs": lds_gbs,
                                  "peak_gbs": LDS_PEAK_GBS, "frac": lds_gbs / LDS_PEAK_GBS,
                                  "peak_note": "128 bytes per clock per CU, 256 CUs, 2.4 GHz"}
    if cnt:
        hbm_cnt = cnt["traffic_bytes_per_launch"] / (launch_ms * 1e-3) / 1e9
        rec["hbm_counters"] = {"achieved_gbs": hbm_cnt, "frac": hbm_cnt / HBM_PEAK_GBS}
    if cnt and cnt.get("valu_insts_per_launch"):
        rate = cnt["valu_insts_per_launch"] / (launch_ms * 1e-3)
        rec.update({"bound": "valu", "achieved": rate, "peak": valu_peak,
                    "unit": "wave64 VALU instructions/s", "frac": rate / valu_peak,
                    "valu_insts_per_launch": cnt["valu_insts_per_launch"]})
    else:
        rec.update({"bound": "hbm", "achieved": hbm_alg, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": hbm_alg / HBM_PEAK_GBS})
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--configs", type=int, default=1024)
    ap.add_argument("--T", type=int, default=128)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--N", type=int, default=33)
    ap.add_argument("--cpu-configs", type=int, default=2,
                    help="CPU baseline: configurations per stream and repetition")
    ap.add_argument("--cpu-slices", type=int, default=128,
                    help="CPU baseline: slices per configuration")
    ap.add_argument("--cpu-reps", type=int, default=3, help="CPU baseline: best of this many")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline streams (0: the host's CPU share for one GPU, 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched uneven-shape measurement")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-proposal measurement")
    ap.add_argument("--no-other-mode", action="store_true",
                    help="skip the run of the workload in the other conjugate-partner mode")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the dense-sink (MFMA overlap) measurement")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the FP64 record at L=16, N=7, T=64, 2048 configurations")
    ap.add_argument("--no-pair", action="store_true",
                    help="skip the pair-sink (SlaterDet3) measurement")
    ap.add_argument("--precision-configs", type=int, default=1024,
                    help="configurations of the FP32-vs-FP64 correlator check (0: skip)")
    ap.add_argument("--conj", type=int, default=1,
                    help="1: conjugate partners (one orbital of each (k, -k) pair evolved); "
                         "0: every orbital evolved, as the reference does")
    ap.add_argument("--acc", type=int, default=1,
                    help="FP32 arithmetic: 1 FP64 x-chain and z-rows (default), 2 FP64 x-chain "
                         "with FP32 z-rows, 0 plain FP32")
    ap.add_argument("--path", type=int, default=0,
                    help="evolution kernels: 0 auto, 1 two passes per slice, 2 single pass")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus N > 1 is not supported",
              file=sys.stderr)
        sys.exit(2)

    import nrcps_amd as na

    L, N, T, C = args.L, args.N, args.T, args.configs
    tau, A = 0.05, 0.4
    V = L ** 3
    ctx = na.Context(0)
    moms = _lowest(L, N)
    kin = _tables(L, tau, 1.0)
    ens = na.Ensemble(ctx, L, moms, kin, A, C, prec=na.PREC_F32, split=na.SPLIT_KSPACE)
    ens.set_path(args.path)
    ens.set_symmetry(bool(args.conj))
    ens.set_accuracy(args.acc)
    seed = 7012

    for w in range(args.warmup):
        ens.run(seed, (1 << 24) + w * C, T)
        print(f"warmup step {w} done", file=sys.stderr, flush=True)

    ctx.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        la, ph, _ = ens.run(seed, k * C, T)
    ctx.sync()
    dt = time.perf_counter() - t0
    ms_per_step = 1e3 * dt / args.steps
    cs_per_s = C * T * args.steps / dt
    assert np.all(np.isfinite(la)), "non-finite log|det|"
    print(f"timed {args.steps} steps: {ms_per_step:.1f} ms/step", file=sys.stderr, flush=True)

    # per-kernel durations from HIP events on the ensemble's stream (one extra, untimed step)
    ens.profile(True)
    ens.run(seed, 1 << 23, T)
    tx, nx = ens.kernel_time(0)
    tyz, nyz = ens.kernel_time(1)
    tlu, nlu = ens.kernel_time(2)
    tf, nf = ens.kernel_time(3)
    tg, ng = ens.kernel_time(4)
    ens.profile(False)
    roofline = evolution_roofline(L, N, T, C, V, ens.volumes_per_config(), args, tx, nx, tyz,
                                  nyz, tlu, nlu, tf, nf, tg, ng)
    ens.close()
    ens = None

    batched = None
    if not args.no_batched:
        batched = batched_uneven(ctx, na, conj=bool(args.conj))
        print(f"batched uneven shapes: {batched['value']:.0f} configs*slices/s",
              file=sys.stderr, flush=True)

    learned = None
    if not args.no_learned:
        learned = learned_proposal(ctx, na, conj=bool(args.conj))
        print(f"fields with memory: gain {learned['memory_fields']['gain_min_over_t']:.2f}x "
              f"effective samples/s", file=sys.stderr, flush=True)
        print(f"memory network: gain {learned['memory_network']['gain_min_over_t']:.2f}x "
              f"effective samples/s", file=sys.stderr, flush=True)
        print(f"fitted tilt: gain {learned['fitted_tilt']['gain_min_over_t']:.2f}x effective "
              f"samples/s", file=sys.stderr, flush=True)
        print(f"learned proposal, 4-slice groups: gain "
              f"{learned['network_grouped']['gain_min_over_t']:.2f}x effective samples/s",
              file=sys.stderr, flush=True)
        print(f"learned proposal (network): gain {learned['network_gain_min_over_t']:.2f}x "
              "effective samples/s",
              file=sys.stderr, flush=True)

    pair = None
    if not args.no_pair:
        pair = pair_sinks(ctx, na, conj=bool(args.conj))
        print(f"pair sinks: {pair['value']:.0f} configs*slices/s", file=sys.stderr, flush=True)

    # the same workload in the other conjugate-partner mode, and in the other arithmetic
    other = None
    if not args.no_other_mode:
        other = {}
        modes = [("conj_partners_" + ("off" if args.conj else "on"), not args.conj, args.acc)]
        modes += [(ACC_KEYS[a], bool(args.conj), a) for a in (0, 1, 2) if a != args.acc]
        for key, conj, acc in modes:
            e = na.Ensemble(ctx, L, moms, kin, A, C, prec=na.PREC_F32, split=na.SPLIT_KSPACE)
            e.set_path(args.path)
            e.set_symmetry(conj)
            e.set_accuracy(acc)
            e.run(seed, (1 << 24), T)
            ctx.sync()
            t0 = time.perf_counter()
            for k in range(args.steps):
                e.run(seed, k * C, T)
            ctx.sync()
            dt2 = time.perf_counter() - t0
            e.profile(True)
            e.run(seed, 1 << 23, T)
            tx2, nx2 = e.kernel_time(0)
            e.profile(False)
            nvo2 = e.volumes_per_config()
            e.close()
            b2 = nvo2 * 16 * V + V
            a2 = b2 * C / (tx2 / T * 1e-3) / 1e9
            other[key] = {"conj_partners": bool(conj), "acc": acc,
                          "value": C * T * args.steps / dt2, "unit": "configs*slices/s",
                          "volumes_per_config": nvo2, "evolution_ms_per_slice": tx2 / T,
                          "hbm_algorithmic": {"achieved_gbs": a2, "frac": a2 / HBM_PEAK_GBS,
                                              "bytes_per_cs": b2}}
            print(f"other mode {key}: {other[key]['value']:.0f} configs*slices/s",
                  file=sys.stderr, flush=True)

    # the FP64 reference of the precision reports (every orbital, same fields)
    ref64 = None
    precision = None
    if args.precision_configs > 0:
        e = na.Ensemble(ctx, L, moms, kin, A, args.precision_configs, prec=na.PREC_F64,
                        split=na.SPLIT_KSPACE)
        ref64 = e.run(seed, 0, T)[:2]
        e.close()
        modes = (bool(args.conj),) if args.no_other_mode else (bool(args.conj), not args.conj)
        reps = precision_report(ctx, na, L, N, T, tau, A, args.precision_configs, seed, modes,
                                acc=args.acc, ref64=ref64)
        precision = reps[0]
        if len(reps) > 1 and other is not None:
            other["conj_partners_" + ("off" if args.conj else "on")]["precision"] = reps[1]
        if other is not None:
            for a in (0, 1, 2):
                if a != args.acc:
                    other[ACC_KEYS[a]]["precision"] = precision_report(
                        ctx, na, L, N, T, tau, A, args.precision_configs, seed,
                        (bool(args.conj),), acc=a, ref64=ref64)[0]
        print(f"precision: max rel err of C(t) {precision['max_rel_err_C']:.2e}",
              file=sys.stderr, flush=True)

    dense = None
    if not args.no_dense:
        dense = dense_overlap(ctx, na, L=L, N=N, T=T, C=C, ref64=ref64 if (
            args.precision_configs == C) else None, acc=args.acc)
        print(f"dense overlaps: {dense['value']:.0f} configs*slices/s with partners "
              f"({dense['every_orbital']['value']:.0f} every orbital), "
              f"{dense['overlap_ms_per_slice']:.3f} ms per slice, MFMA frac "
              f"{dense['mfma']['frac']:.3f}", file=sys.stderr, flush=True)

    fp64c2 = None
    if not args.no_fp64:
        fp64c2 = fp64_config2(ctx, na, threads=0 if args.no_cpu else (
            args.cpu_threads or min(16, os.cpu_count() or 1)))
        print(f"FP64 L=16 N=7: {fp64c2['value']:.0f} configs*slices/s "
              f"({fp64c2['value_conj_partners']:.0f} with partners), log|det| within "
              f"{fp64c2['max_abs_diff_logdet_vs_oracle_first_2_configs']:.1e} of the oracle",
              file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(f"cpu baseline: {threads} streams x {args.cpu_configs} configs x "
              f"{args.cpu_slices} slices, best of {args.cpu_reps}", file=sys.stderr, flush=True)
        cpu = cpu_baseline(L, N, tau, A, args.cpu_configs, args.cpu_slices, threads,
                           args.cpu_reps)

    out = {
        "metric": "aux-field configs*time-slices/sec at L=32,N=33",
        "value": cs_per_s, "unit": "configs*slices/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"Slater-determinant correlator over aux-field configurations: "
                               f"L={L}, N={N}, T={T}, tau={tau}, A={A}, {C} configs per step, "
                               f"evolution in {ACC_NAMES[args.acc]}, FP64 LU",
                   "L": L, "N": N, "T": T, "configs": C,
                   "parallelism": "single GPU, batched configurations"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "precision": precision,
        "batched_uneven": batched,
        "learned_proposal": learned,
        "pair_sinks": pair,
        "other_mode": other,
        "dense_overlap": dense,
        "fp64_config2": fp64c2,
    }
    if cpu:
        out["speedup_vs_cpu"] = cs_per_s / cpu["value"]
        out["speedup_vs_cpu_all_visible_cpus_linear"] = (
            cs_per_s / cpu["value_all_visible_cpus_linear"])
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
