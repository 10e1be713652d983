"""nrcps_amd -- MI355X (gfx950) implementation of nrCPS's auxiliary-field propagator and
Slater-determinant correlator path.

This module is a thin ctypes binding of the C ABI in include/nrcps.h (libnrcps_hip.so).
It never falls back to a CPU path: if the HIP library is missing or no GPU is present,
`Context()` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# NRC_CHECK=1 in the environment selects the check build (nrc_check_build): the same ABI with
# device synchronisation after every launch, index checks and NaN scans of results
LIB_PATH = os.path.join(PKG, "libnrcps_hip_check.so" if os.environ.get("NRC_CHECK") == "1"
                        else "libnrcps_hip.so")

OK = 0
PREC_F32, PREC_F64 = 0, 1
SPLIT_KSPACE, SPLIT_XSPACE = 0, 1
OVERLAP_F64, OVERLAP_F32, OVERLAP_BF16X3, OVERLAP_BF16X6 = 0, 1, 2, 3
PATH_AUTO, PATH_TWO_PASS, PATH_SINGLE_PASS = 0, 1, 2
ACC_FP32, ACC_X64, ACC_XCHAIN = 0, 1, 2
OVL_AUTO, OVL_COLS, OVL_SHARED, OVL_SHARED_DEEP, OVL_SHARED_PIPE = 0, 1, 2, 3, 4
SUPPORTED_L = (4, 6, 8, 12, 16, 24, 32)

# every symbol include/nrcps.h declares
EXPORTS = [
    "nrc_init", "nrc_finalize", "nrc_error_string", "nrc_device_name", "nrc_malloc", "nrc_free",
    "nrc_memcpy_h2d", "nrc_memcpy_d2h", "nrc_memcpy_d2d", "nrc_memset", "nrc_sync",
    "nrc_fft_plan", "nrc_fft_exec", "nrc_fft_destroy", "nrc_scale_real", "nrc_mul_complex",
    "nrc_plus_one_div_vol",
    "nrc_overlap", "nrc_pair_overlap", "nrc_lu_det", "nrc_ludcmp", "nrc_ref_abs", "nrc_z2_bits", "nrc_ensemble_create", "nrc_ensemble_run",
    "nrc_ensemble_run_fields", "nrc_ensemble_orbitals", "nrc_ensemble_set_path",
    "nrc_ensemble_set_symmetry", "nrc_ensemble_volumes", "nrc_ensemble_set_trap",
    "nrc_ensemble_set_accuracy", "nrc_batch_captures", "nrc_overlap_noalloc", "nrc_z2_sums",
    "nrc_ensemble_set_tilt", "nrc_ensemble_tilt", "nrc_ensemble_log_ratio",
    "nrc_ensemble_set_memory", "nrc_ensemble_memory", "nrc_z2_memory_eval",
    "nrc_ensemble_set_memnet", "nrc_z2_memnet_eval", "nrc_z2_memnet_grad",
    "nrc_overlap_set_variant", "nrc_overlap_variant", "nrc_overlap_tune",
    "nrc_ensemble_tune_report", "nrc_ensemble_launch_slices", "nrc_tune_save", "nrc_tune_load",
    "nrc_last_error_detail", "nrc_check_build", "nrc_ensemble_evolve_fused",
    "nrc_scale_const", "nrc_norms",
    "nrc_ensemble_profile",
    "nrc_ensemble_kernel_time", "nrc_ensemble_results", "nrc_ensemble_tune",
    "nrc_ensemble_destroy", "nrc_batch_create", "nrc_batch_run", "nrc_batch_capture_ms",
    "nrc_batch_destroy", "nrc_ensemble_begin", "nrc_ensemble_measure", "nrc_ensemble_evolve",
    "nrc_ensemble_results_dev", "nrc_proposal_create", "nrc_proposal_destroy",
    "nrc_proposal_param_count", "nrc_proposal_set_params", "nrc_proposal_get_params",
    "nrc_proposal_forward", "nrc_proposal_sample", "nrc_proposal_train_step", "nrc_norm_logw",
    "nrc_z2_bits_list",
]


class NrcError(RuntimeError):
    pass


class EnsembleDesc(C.Structure):
    _fields_ = [("L", C.c_int), ("n_orb", C.c_int), ("n_configs", C.c_int), ("prec", C.c_int),
                ("split", C.c_int), ("minors", C.c_int), ("moms", C.c_void_p),
                ("kin", C.c_void_p), ("A", C.c_double), ("smear", C.c_void_p),
                ("sinks", C.c_void_p), ("pair_w", C.c_void_p), ("pair_opp", C.c_void_p)]


_lib = None


def load_library() -> C.CDLL:
    """Loads libnrcps_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NrcError(f"{LIB_PATH} is missing: run `python -m nrcps_amd.build` first")
    L = C.CDLL(LIB_PATH)
    vp, sz, i, u32, u64, dp = C.c_void_p, C.c_size_t, C.c_int, C.c_uint32, C.c_uint64, C.c_void_p
    sig = {
        "nrc_init": [i, C.POINTER(vp)], "nrc_finalize": [vp], "nrc_error_string": [i],
        "nrc_last_error_detail": [], "nrc_check_build": [],
        "nrc_ensemble_evolve_fused": [vp, i, i, vp],
        "nrc_device_name": [vp, C.c_char_p, i, C.POINTER(i)],
        "nrc_malloc": [vp, C.POINTER(vp), sz], "nrc_free": [vp, vp],
        "nrc_memcpy_h2d": [vp, vp, vp, sz], "nrc_memcpy_d2h": [vp, vp, vp, sz],
        "nrc_memcpy_d2d": [vp, vp, vp, sz], "nrc_memset": [vp, vp, i, sz], "nrc_sync": [vp],
        "nrc_fft_plan": [vp, i, i, i, i, i, C.POINTER(vp)], "nrc_fft_exec": [vp, vp, i],
        "nrc_fft_destroy": [vp],
        "nrc_scale_real": [vp, vp, vp, i, C.c_long, i, i],
        "nrc_mul_complex": [vp, vp, vp, i, C.c_long, i],
        "nrc_plus_one_div_vol": [vp, vp, C.c_long, i],
        "nrc_overlap": [vp, vp, vp, vp, i, i, C.c_long, i, i],
        "nrc_overlap_noalloc": [vp, vp, vp, vp, i, i, C.c_long, i, i],
        "nrc_overlap_set_variant": [vp, i, i, C.c_long, i, i],
        "nrc_overlap_variant": [vp, i, i, C.c_long, i, C.POINTER(i)],
        "nrc_overlap_tune": [vp, vp, vp, vp, i, i, C.c_long, i, i, i, C.POINTER(i), dp],
        "nrc_pair_overlap": [vp, vp, vp, vp, vp, vp, i, i, C.c_long, i, i],
        "nrc_lu_det": [vp, vp, i, i, i, i, vp, vp, vp, vp],
        "nrc_ludcmp": [vp, vp, i, i, vp, vp, vp], "nrc_ref_abs": [vp, vp, vp, C.c_long],
        "nrc_z2_bits": [vp, u64, u32, i, u32, u32, i, vp],
        "nrc_ensemble_create": [vp, C.POINTER(EnsembleDesc), C.POINTER(vp)],
        "nrc_ensemble_run": [vp, u64, u32, i, dp, dp, dp],
        "nrc_ensemble_run_fields": [vp, vp, i, dp, dp, dp],
        "nrc_ensemble_orbitals": [vp, i, vp], "nrc_ensemble_set_path": [vp, i],
        "nrc_ensemble_set_symmetry": [vp, i], "nrc_ensemble_volumes": [vp, C.POINTER(i)],
        "nrc_ensemble_set_trap": [vp, vp], "nrc_ensemble_set_accuracy": [vp, i],
        "nrc_ensemble_set_tilt": [vp, C.c_double],
        "nrc_ensemble_set_memory": [vp, C.c_double, C.c_double, C.c_double],
        "nrc_ensemble_memory": [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                C.POINTER(C.c_double)],
        "nrc_z2_memory_eval": [vp, u64, u32, i, i, i, C.c_double, C.c_double, C.c_double, vp],
        "nrc_ensemble_set_memnet": [vp, vp],
        "nrc_z2_memnet_eval": [vp, u64, u32, i, i, i, vp, vp],
        "nrc_z2_memnet_grad": [vp, u64, u32, i, i, i, vp, vp, vp],
        "nrc_ensemble_tilt": [vp, C.POINTER(C.c_double)], "nrc_ensemble_log_ratio": [vp, i, dp],
        "nrc_scale_const": [vp, vp, C.c_long, C.c_double, i, i],
        "nrc_norms": [vp, vp, i, C.c_long, i, vp],
        "nrc_ensemble_profile": [vp, i],
        "nrc_ensemble_kernel_time": [vp, i, C.POINTER(C.c_double), C.POINTER(C.c_long)],
        "nrc_ensemble_results": [vp, dp, dp, dp],
        "nrc_ensemble_tune": [vp, i, C.POINTER(i), C.POINTER(C.c_double)],
        "nrc_ensemble_tune_report": [vp, i, C.POINTER(i), C.POINTER(i), C.POINTER(C.c_double)],
        "nrc_ensemble_launch_slices": [vp, C.POINTER(i)],
        "nrc_tune_save": [vp, C.c_char_p], "nrc_tune_load": [vp, C.c_char_p],
        "nrc_ensemble_destroy": [vp],
        "nrc_batch_create": [vp, vp, i, C.POINTER(vp)],
        "nrc_batch_run": [vp, u64, vp, i, i],
        "nrc_batch_capture_ms": [vp, C.POINTER(C.c_double)],
        "nrc_batch_captures": [vp, C.POINTER(C.c_long)],
        "nrc_batch_destroy": [vp],
        "nrc_ensemble_begin": [vp, i], "nrc_ensemble_measure": [vp, i],
        "nrc_ensemble_evolve": [vp, i, vp],
        "nrc_ensemble_results_dev": [vp, C.POINTER(vp), C.POINTER(vp)],
        "nrc_proposal_create": [vp, i, i, C.POINTER(vp)], "nrc_proposal_destroy": [vp],
        "nrc_proposal_param_count": [], "nrc_proposal_set_params": [vp, vp],
        "nrc_proposal_get_params": [vp, vp], "nrc_proposal_forward": [vp, vp, vp, i, vp],
        "nrc_proposal_sample": [vp, vp, vp, i, u64, u32, u32, u32, vp, vp],
        "nrc_proposal_train_step": [vp, vp, vp, vp, vp, i, C.c_double, C.c_float, i, vp,
                                    C.POINTER(C.c_double)],
        "nrc_norm_logw": [vp, vp, i, i, C.c_double, C.c_double, vp],
        "nrc_z2_bits_list": [vp, u64, vp, vp, i, u32, i, vp],
        "nrc_z2_sums": [vp, u64, u32, i, i, u32, i, vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = (C.c_char_p if name in ("nrc_error_string", "nrc_last_error_detail")
                     else C.c_int)
    _lib = L
    return L


def _check(rc: int, what: str) -> None:
    if rc != OK:
        lib = load_library()
        msg = lib.nrc_error_string(rc).decode()
        detail = lib.nrc_last_error_detail().decode()
        raise NrcError(f"{what} failed: {msg} ({rc})" + (f" [{detail}]" if detail else ""))


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


class DeviceArray:
    """A device allocation with a shape/dtype, owned by a Context."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx, self.shape, self.dtype = ctx, tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(ctx.lib.nrc_malloc(ctx.h, C.byref(p), self.nbytes), "nrc_malloc")
        self.ptr = p

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        _check(self.ctx.lib.nrc_memcpy_h2d(self.ctx.h, self.ptr, _ptr(a), a.nbytes), "h2d")
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _check(self.ctx.lib.nrc_memcpy_d2h(self.ctx.h, _ptr(out), self.ptr, self.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr and getattr(self.ctx, "h", None):
            self.ctx.lib.nrc_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One GPU: a HIP stream and the twiddle tables (nrc_init)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib.nrc_init(device, C.byref(h)), "nrc_init")
        self.h = h
        buf = C.create_string_buffer(256)
        ncu = C.c_int()
        self.lib.nrc_device_name(self.h, buf, 256, C.byref(ncu))
        self.arch, self.n_cu = buf.value.decode(), ncu.value

    def array(self, shape, dtype) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def to_device(self, a: np.ndarray) -> DeviceArray:
        return DeviceArray(self, a.shape, a.dtype).upload(a)

    def sync(self):
        _check(self.lib.nrc_sync(self.h), "nrc_sync")

    # ---------------------------------------------------------------- operators
    def fft3d(self, buf: DeviceArray, L, batch: int, sign: int, prec: int):
        """In-place 3D DFT of `batch` volumes; L is the side of a cubic lattice or (X, Y, Z)."""
        lx, ly, lz = (L, L, L) if np.isscalar(L) else tuple(int(v) for v in L)
        plan = C.c_void_p()
        _check(self.lib.nrc_fft_plan(self.h, lx, ly, lz, batch, prec, C.byref(plan)),
               "fft_plan")
        try:
            _check(self.lib.nrc_fft_exec(plan, buf.ptr, sign), "fft_exec")
        finally:
            self.lib.nrc_fft_destroy(plan)

    def scale_real(self, buf, table, batch, vol, op, prec):
        _check(self.lib.nrc_scale_real(self.h, buf.ptr, table.ptr, batch, vol, op, prec),
               "scale_real")

    def mul_complex(self, buf, field, batch, vol, prec):
        _check(self.lib.nrc_mul_complex(self.h, buf.ptr, field.ptr, batch, vol, prec),
               "mul_complex")

    def overlap(self, phi, psi, M, ns, nt, vol, ncfg, mode):
        _check(self.lib.nrc_overlap(self.h, phi.ptr, psi.ptr, M.ptr, ns, nt, vol, ncfg, mode),
               "overlap")

    def tune_save(self, path: str):
        """The context's tuned choices to a text file (nrc_tune_save)."""
        _check(self.lib.nrc_tune_save(self.h, os.fsencode(path)), "tune_save")

    def tune_load(self, path: str):
        _check(self.lib.nrc_tune_load(self.h, os.fsencode(path)), "tune_load")

    def set_overlap_variant(self, ns, nt, vol, mode, variant):
        """Pins the split-BF16 GEMM variant (OVL_*) for a shape; OVL_AUTO clears the pin."""
        _check(self.lib.nrc_overlap_set_variant(self.h, ns, nt, vol, mode, variant),
               "overlap_set_variant")

    def overlap_variant(self, ns, nt, vol, mode) -> int:
        v = C.c_int()
        _check(self.lib.nrc_overlap_variant(self.h, ns, nt, vol, mode, C.byref(v)),
               "overlap_variant")
        return v.value

    def overlap_tune(self, phi, psi, M, ns, nt, vol, ncfg, mode, reps=3):
        """Times every GEMM variant, keeps the fastest: (best, {variant: ms})."""
        best, ms = C.c_int(), np.zeros(4)
        _check(self.lib.nrc_overlap_tune(self.h, phi.ptr, psi.ptr, M.ptr, ns, nt, vol, ncfg,
                                         mode, reps, C.byref(best), _ptr(ms)), "overlap_tune")
        return best.value, dict(zip((OVL_COLS, OVL_SHARED, OVL_SHARED_DEEP, OVL_SHARED_PIPE),
                                    ms.tolist()))

    def scale_const(self, buf, n, d, op, prec):
        """nrc_scale_const: buf *= d (op 0) or /= d (op 1) over n complex values."""
        _check(self.lib.nrc_scale_const(self.h, buf.ptr, n, d, op, prec), "scale_const")

    def norms(self, buf, batch, vol, prec):
        """nrc_norms: the L2 norm of each of `batch` volumes of `vol` complex values."""
        out = np.zeros(batch)
        _check(self.lib.nrc_norms(self.h, buf.ptr, batch, vol, prec, out.ctypes.data), "norms")
        return out

    def pair_overlap(self, psi1, psi2, w, opp, M, n1, n2, vol, ncfg, mode):
        """nrc_pair_overlap: M[c][s][t] = sum_i w[i] psi1[c][s][i] psi2[c][t][opp[i]]."""
        _check(self.lib.nrc_pair_overlap(self.h, psi1.ptr, psi2.ptr, w.ptr, opp.ptr, M.ptr, n1,
                                         n2, vol, ncfg, mode), "pair_overlap")

    def lu_det(self, M, n, ncfg, minors=False, ldm=None):
        nm = n if minors else 1
        la = self.array((ncfg, nm), np.float64)
        ph = self.array((ncfg, nm), np.complex128)
        ag = self.array((ncfg, nm), np.float64)
        st = self.array((ncfg, nm), np.int32)
        _check(self.lib.nrc_lu_det(self.h, M.ptr, n, ldm or n, ncfg, int(minors), la.ptr,
                                   ph.ptr, ag.ptr, st.ptr), "lu_det")
        return la.download(), ph.download(), ag.download(), st.download()

    def ludcmp(self, a: np.ndarray):
        """Batched ludcmp of contiguous complex matrices a[c][n][n]: (lu, indx, d, status)."""
        ncfg, n, _ = a.shape
        d_a = self.to_device(np.ascontiguousarray(a, np.complex128))
        indx = self.array((ncfg, n), np.int32)
        d = self.array((ncfg,), np.float64)
        st = self.array((ncfg,), np.int32)
        _check(self.lib.nrc_ludcmp(self.h, d_a.ptr, n, ncfg, indx.ptr, d.ptr, st.ptr), "ludcmp")
        return d_a.download(), indx.download(), d.download(), st.download()

    def ref_abs(self, z: np.ndarray) -> np.ndarray:
        z = np.ascontiguousarray(z, np.complex128).reshape(-1)
        out = self.array(z.shape, np.float64)
        _check(self.lib.nrc_ref_abs(self.h, self.to_device(z).ptr, out.ptr, z.size), "ref_abs")
        return out.download()

    def z2_bits(self, seed, config0, nconf, t, vol, stream=0) -> np.ndarray:
        nw = (vol + 31) // 32
        w = self.array((nconf, nw), np.uint32)
        _check(self.lib.nrc_z2_bits(self.h, seed, config0, nconf, t, stream, vol, w.ptr),
               "z2_bits")
        return w.download()

    def close(self):
        if getattr(self, "h", None):
            self.lib.nrc_finalize(self.h)
            self.h = None


class Ensemble:
    """The batched production slice loop (nrc_ensemble_*)."""

    def __init__(self, ctx: Context, L, moms, kin, A, n_configs, prec=PREC_F32,
                 split=SPLIT_KSPACE, minors=False, smear=None, sinks=None, pair_w=None,
                 pair_opp=None):
        """pair_w / pair_opp: the pair-sink measurement (2N results per slice, SlaterDet3's:
        m = 2n + 1 leading (n+1) minor of the pair matrix, m = 2n with column n from the
        one-body overlaps)."""
        self.ctx, self.lib = ctx, ctx.lib
        self.L, self.N, self.C, self.prec = L, len(moms), n_configs, prec
        self.nm = 2 * self.N if pair_w is not None else (self.N if minors else 1)
        self._moms = np.ascontiguousarray(moms, dtype=np.int32).reshape(-1)
        self._kin = np.ascontiguousarray(kin, dtype=np.float64)
        d = EnsembleDesc()
        d.L, d.n_orb, d.n_configs, d.prec, d.split, d.minors = (L, self.N, n_configs, prec, split,
                                                               int(minors))
        d.moms, d.kin, d.A = self._moms.ctypes.data, self._kin.ctypes.data, A
        self._smear = None
        if smear is not None:
            self._smear = np.ascontiguousarray(smear, dtype=np.float64)
            d.smear = self._smear.ctypes.data
        self._sinks = None
        if sinks is not None:
            self._sinks = np.ascontiguousarray(sinks, dtype=np.complex128).reshape(-1)
            d.sinks = self._sinks.ctypes.data
        if pair_w is not None:
            self._pair = (np.ascontiguousarray(pair_w, dtype=np.float64),
                          np.ascontiguousarray(pair_opp, dtype=np.int32))
            d.pair_w, d.pair_opp = self._pair[0].ctypes.data, self._pair[1].ctypes.data
        h = C.c_void_p()
        _check(self.lib.nrc_ensemble_create(ctx.h, C.byref(d), C.byref(h)), "ensemble_create")
        self.h = h

    def _outs(self, T):
        return (np.zeros((T, self.C, self.nm)), np.zeros((T, self.C, self.nm), np.complex128),
                np.zeros((T, self.C, self.nm)))

    def run(self, seed: int, config0: int, T: int):
        la, ph, ag = self._outs(T)
        _check(self.lib.nrc_ensemble_run(self.h, seed, config0, T, _ptr(la), _ptr(ph), _ptr(ag)),
               "ensemble_run")
        return la, ph, ag

    # slice-by-slice stepping (nrc_ensemble_begin / measure / evolve)
    def begin(self, T: int):
        _check(self.lib.nrc_ensemble_begin(self.h, T), "ensemble_begin")

    def measure(self, t: int):
        _check(self.lib.nrc_ensemble_measure(self.h, t), "ensemble_measure")

    def evolve(self, t: int, words: "DeviceArray"):
        _check(self.lib.nrc_ensemble_evolve(self.h, t, words.ptr), "ensemble_evolve")

    def evolve_fused(self, t: int, S: int, words: "DeviceArray"):
        """Slices t .. t+S-1 in one launch with fields words [S][C][nw], rows t .. t+S-1
        measured inside (nrc_ensemble_evolve_fused)."""
        _check(self.lib.nrc_ensemble_evolve_fused(self.h, t, S, words.ptr),
               "ensemble_evolve_fused")

    def results_dev(self):
        """Device addresses (ints) of the [T][C][m] log|det| and phase result rows."""
        la, ph = C.c_void_p(), C.c_void_p()
        _check(self.lib.nrc_ensemble_results_dev(self.h, C.byref(la), C.byref(ph)),
               "ensemble_results_dev")
        return la.value, ph.value

    def run_fields(self, words: np.ndarray, T: int):
        w = np.ascontiguousarray(words, dtype=np.uint32)
        la, ph, ag = self._outs(T)
        _check(self.lib.nrc_ensemble_run_fields(self.h, _ptr(w), T, _ptr(la), _ptr(ph),
                                                _ptr(ag)), "ensemble_run_fields")
        return la, ph, ag

    def orbitals(self, c: int) -> np.ndarray:
        dt = np.complex64 if self.prec == PREC_F32 else np.complex128
        out = np.zeros((self.N, self.L ** 3), dtype=dt)
        _check(self.lib.nrc_ensemble_orbitals(self.h, c, _ptr(out)), "ensemble_orbitals")
        return out

    def set_path(self, path: int):
        """PATH_AUTO, PATH_TWO_PASS or PATH_SINGLE_PASS (nrc_ensemble_set_path)."""
        _check(self.lib.nrc_ensemble_set_path(self.h, path), "ensemble_set_path")

    def set_accuracy(self, mode: int):
        """ACC_FP32, ACC_X64 (the default) or ACC_XCHAIN for FP32 ensembles
        (nrc_ensemble_set_accuracy)."""
        _check(self.lib.nrc_ensemble_set_accuracy(self.h, mode), "ensemble_set_accuracy")

    def set_tilt(self, p_plus: float) -> float:
        """Tilted fields (nrc_ensemble_set_tilt): each site +1 with probability p_plus (0.5:
        the plain sampler); returns the exact probability in use (a multiple of 2^-32)."""
        _check(self.lib.nrc_ensemble_set_tilt(self.h, float(p_plus)), "ensemble_set_tilt")
        q = C.c_double()
        _check(self.lib.nrc_ensemble_tilt(self.h, C.byref(q)), "ensemble_tilt")
        return q.value

    def set_memory(self, lam: float, theta: float, rho: float):
        """Fields with memory (nrc_ensemble_set_memory): each site +1 with probability
        (1 + lam + theta m) / 2, m = sum_k rho^(k-1) phi(t - k) its own decayed past field
        (lam = theta = 0: the plain sampler); returns the realised (lam, theta, rho)."""
        _check(self.lib.nrc_ensemble_set_memory(self.h, float(lam), float(theta), float(rho)),
               "ensemble_set_memory")
        return self.memory()

    def set_memnet(self, params):
        """The memory network (nrc_ensemble_set_memnet): 35 FP32 parameters, or None for the
        plain sampler."""
        if params is None:
            _check(self.lib.nrc_ensemble_set_memnet(self.h, None), "ensemble_set_memnet")
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.shape != (35,):
            raise ValueError("35 parameters")
        _check(self.lib.nrc_ensemble_set_memnet(self.h, p.ctypes.data), "ensemble_set_memnet")

    def memory(self):
        out = [C.c_double() for _ in range(3)]
        _check(self.lib.nrc_ensemble_memory(self.h, *[C.byref(x) for x in out]),
               "ensemble_memory")
        return tuple(x.value for x in out)

    def log_ratio(self, T: int) -> np.ndarray:
        """[T][C] log of the plain-to-proposal density ratio of the last run's fields before
        each measured slice (tilted fields or fields with memory; zero for a plain run); T is
        the last run's slice count."""
        out = np.zeros((T, self.C), dtype=np.float64)
        _check(self.lib.nrc_ensemble_log_ratio(self.h, T, _ptr(out)), "ensemble_log_ratio")
        return out

    def set_symmetry(self, conj: bool):
        """Conjugate partners (nrc_ensemble_set_symmetry): evolve one orbital of each (k, -k)
        pair and read the other's sink values as the conjugate mirror.  Call before a run."""
        _check(self.lib.nrc_ensemble_set_symmetry(self.h, int(conj)), "ensemble_set_symmetry")

    def set_trap(self, q):
        """x-space multiplier q (V reals) applied with the field every slice, e.g. (V p)^2 for
        the trapped order K, P, I, P, K with Potential's table p; None removes it
        (nrc_ensemble_set_trap)."""
        if q is None:
            _check(self.lib.nrc_ensemble_set_trap(self.h, None), "ensemble_set_trap")
            return
        qa = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        assert qa.size == self.L ** 3
        _check(self.lib.nrc_ensemble_set_trap(self.h, qa.ctypes.data), "ensemble_set_trap")

    def volumes_per_config(self) -> int:
        n = C.c_int()
        _check(self.lib.nrc_ensemble_volumes(self.h, C.byref(n)), "ensemble_volumes")
        return n.value

    def profile(self, enable: bool):
        _check(self.lib.nrc_ensemble_profile(self.h, int(enable)), "ensemble_profile")

    def kernel_time(self, kind: int):
        ms, n = C.c_double(), C.c_long()
        _check(self.lib.nrc_ensemble_kernel_time(self.h, kind, C.byref(ms), C.byref(n)),
               "kernel_time")
        return ms.value, n.value

    def results(self, T: int):
        """The last run's [t][c][m] outputs (also after a batch run)."""
        la, ph, ag = self._outs(T)
        _check(self.lib.nrc_ensemble_results(self.h, _ptr(la), _ptr(ph), _ptr(ag)),
               "ensemble_results")
        return la, ph, ag

    def tune(self, T_probe: int = 4):
        """Autotuner (nrc_ensemble_tune): returns (chosen path, ms per slice)."""
        p, ms = C.c_int(), C.c_double()
        _check(self.lib.nrc_ensemble_tune(self.h, T_probe, C.byref(p), C.byref(ms)),
               "ensemble_tune")
        return p.value, ms.value

    def tune_report(self):
        """The last tune's timings: [(kind, value, ms)], kind 'path', 'slices', 'gemm' or
        'tile'."""
        out, i = [], 0
        k, v, ms = C.c_int(), C.c_int(), C.c_double()
        names = {0: "path", 1: "slices", 2: "gemm", 3: "tile"}
        while self.lib.nrc_ensemble_tune_report(self.h, i, C.byref(k), C.byref(v),
                                                C.byref(ms)) == OK:
            out.append((names[k.value], v.value, ms.value))
            i += 1
        return out

    def launch_slices(self) -> int:
        n = C.c_int()
        _check(self.lib.nrc_ensemble_launch_slices(self.h, C.byref(n)), "ensemble_launch_slices")
        return n.value

    def close(self):
        # an ensemble lives on its context's device; after nrc_finalize it is gone with it.
        # While it belongs to a batch the destroy is refused (NrcError) and the handle kept.
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            _check(self.lib.nrc_ensemble_destroy(self.h), "ensemble_destroy")
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Shape groups evolved together (nrc_batch_*): one Ensemble per (L, N) group."""

    def __init__(self, ctx: Context, groups):
        self.ctx, self.lib, self.groups = ctx, ctx.lib, list(groups)
        arr = (C.c_void_p * len(self.groups))(*[g.h for g in self.groups])
        h = C.c_void_p()
        _check(self.lib.nrc_batch_create(ctx.h, arr, len(self.groups), C.byref(h)),
               "batch_create")
        self.h = h

    def run(self, seed: int, config0, T: int, graph: bool = True):
        c0 = np.ascontiguousarray(config0, dtype=np.uint32)
        assert c0.size == len(self.groups)
        _check(self.lib.nrc_batch_run(self.h, seed, _ptr(c0), T, int(graph)), "batch_run")

    def capture_ms(self) -> float:
        ms = C.c_double()
        _check(self.lib.nrc_batch_capture_ms(self.h, C.byref(ms)), "batch_capture_ms")
        return ms.value

    def captures(self) -> int:
        n = C.c_long()
        _check(self.lib.nrc_batch_captures(self.h, C.byref(n)), "batch_captures")
        return n.value

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.lib.nrc_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
