"""ctypes bindings for the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is the parity checker for the HIP product path;
only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LUDCMP_PATH = os.path.join(HERE, "_ref", "libref_ludcmp.so")

DISP_STANDARD, DISP_QUADRATIC, DISP_PERFECT = 0, 1, 2
CUTOFF_NONE, CUTOFF_HARD = 0, 1
INTER_NONE, INTER_XISQMINUSONE, INTER_ONEMINUSXIINVSQ = 0, 1, 2
SPLIT_KSPACE, SPLIT_XSPACE = 0, 1
FIELD_PHILOX, FIELD_RANLXS, FIELD_ARRAY, FIELD_REAL = 0, 1, 2, 3
WAVEFUNC_NONE, WAVEFUNC_UNIFORM, WAVEFUNC_GND, WAVEFUNC_TGND, WAVEFUNC_PAIR1, WAVEFUNC_PAIR2 = range(6)

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class Ranlxs(C.Structure):
    _fields_ = [("x", C.c_int * 96), ("carry", C.c_int * 4), ("pr", C.c_int),
                ("prm", C.c_int), ("ir", C.c_int), ("jr", C.c_int), ("is_", C.c_int),
                ("is_old", C.c_int), ("level", C.c_int)]


class RunDesc(C.Structure):
    _fields_ = [("L", C.c_int), ("N", C.c_int), ("T", C.c_int), ("split", C.c_int),
                ("minors", C.c_int), ("moms", C.c_void_p), ("kin", C.c_void_p),
                ("A", C.c_double), ("smear", C.c_void_p), ("field_kind", C.c_int),
                ("seed", C.c_uint64), ("config", C.c_uint32), ("rng", C.c_void_p),
                ("phi", C.c_void_p), ("pair_w", C.c_void_p), ("pair_opp", C.c_void_p),
                ("sources", C.c_void_p), ("pot", C.c_void_p), ("phi_real", C.c_void_p),
                ("sinks", C.c_void_p), ("norms", C.c_void_p), ("gather", C.c_int)]


def build(force: bool = False) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_fft3d.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_fft1d.argtypes = [_dp, C.c_int, C.c_int, C.c_int]
        L.orc_dispersion.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _ip, C.c_int,
                                     C.c_double, C.c_int, C.c_double]
        L.orc_dispersion.restype = None
        L.orc_kinetic_xi.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _ip, C.c_int,
                                     C.c_double, C.c_double, C.c_double]
        L.orc_kinetic_xi.restype = None
        L.orc_kinetic_exp.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
        L.orc_kinetic_exp.restype = None
        L.orc_potential_table.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_potential_table.restype = C.c_int
        L.orc_interaction_table.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _ip, C.c_double,
                                            C.c_int, _dp, C.c_int, C.c_double, C.c_double]
        L.orc_interaction_field.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int]
        L.orc_momenta.argtypes = [C.c_int, C.c_int, C.c_int, _ip, C.c_int, C.c_double,
                                  C.c_double, C.c_double, _ip, _ip, _dp, C.c_int]
        L.orc_lowest_momenta.argtypes = [C.c_int, C.c_int, _ip]
        L.orc_ranlxs_init.argtypes = [C.POINTER(Ranlxs), C.c_int, C.c_long]
        L.orc_ranlxs_next.argtypes = [C.POINTER(Ranlxs)]
        L.orc_ranlxs_next.restype = C.c_float
        L.orc_ranlxs_get_state.argtypes = [C.POINTER(Ranlxs), _lp]
        L.orc_ranlxs_get_state.restype = None
        L.orc_ranlxs_set_state.argtypes = [C.POINTER(Ranlxs), _lp]
        L.orc_ranlxs_z2_field.argtypes = [C.POINTER(Ranlxs), C.c_int, _i8p]
        L.orc_ranlxs_z2_field.restype = None
        L.orc_gauss.argtypes = [C.POINTER(Ranlxs), C.c_double]
        L.orc_gauss.restype = C.c_double
        L.orc_gauss_field.argtypes = [C.POINTER(Ranlxs), C.c_int, C.c_double, _dp]
        L.orc_gauss_field.restype = None
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_philox4x32_10.restype = None
        L.orc_z2_bits.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, _u32p]
        L.orc_z2_bits.restype = None
        L.orc_bits_to_phi.argtypes = [_u32p, C.c_int, _i8p]
        L.orc_bits_to_phi.restype = None
        L.orc_ludcmp.argtypes = [_dp, C.c_int, _ip, C.POINTER(C.c_double)]
        L.orc_lu_det.argtypes = [_dp, C.c_int, C.c_double, C.POINTER(C.c_double), _dp,
                                 C.POINTER(C.c_double), _dp]
        L.orc_lu_det.restype = None
        L.orc_overlap.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_long]
        L.orc_overlap.restype = None
        L.orc_run_config.argtypes = [C.POINTER(RunDesc), C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p]
        L.orc_exact_c1.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, _dp]
        L.orc_dawson.argtypes = [C.c_double]
        L.orc_dawson.restype = C.c_double
        L.orc_opposite_parity_index.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip]
        L.orc_opposite_parity_index.restype = None
        L.orc_two_body_wavefunc.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _ip, C.c_int,
                                            C.c_int, C.c_int, C.c_double, C.c_double,
                                            C.c_double, C.c_double]
        L.orc_two_body_wavefunc.restype = None
        L.orc_pair_overlap.argtypes = [_dp, _dp, _dp, _ip, _dp, C.c_int, C.c_int, C.c_long]
        L.orc_pair_overlap.restype = None
        L.orc_hermite.argtypes = [_dp, C.c_int, C.c_double]
        L.orc_hermite.restype = None
        L.orc_sho_source.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_sho_source.restype = None
        L.orc_time_threads.argtypes = [C.POINTER(RunDesc), C.c_int, C.c_int]
        L.orc_time_threads.restype = C.c_double
        L.orc_time_pool.argtypes = [C.POINTER(RunDesc), C.c_int, C.c_int, C.c_int, C.c_int,
                                    _dp]
        L.orc_time_pool.restype = C.c_int
        _lib = L
    return _lib


# ----------------------------------------------------------------------------------------
# numpy-level helpers
# ----------------------------------------------------------------------------------------
def _bc(bc):
    return np.ascontiguousarray(bc if bc is not None else [0, 0, 0], dtype=np.int32)


def fft3d(a: np.ndarray, sign: int) -> np.ndarray:
    """Unnormalised 3D DFT over the last three axes of a complex array (copy)."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    shp = a.shape
    lx, ly, lz = shp[-3:]
    batch = int(np.prod(shp[:-3])) if len(shp) > 3 else 1
    buf = a.reshape(-1).view(np.float64).copy()
    rc = lib().orc_fft3d(buf, lx, ly, lz, batch, sign)
    assert rc == 0
    return buf.view(np.complex128).reshape(shp)


def fft1d(a: np.ndarray, sign: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.complex128)
    n = a.shape[-1]
    buf = a.reshape(-1).view(np.float64).copy()
    assert lib().orc_fft1d(buf, n, a.size // n, sign) == 0
    return buf.view(np.complex128).reshape(a.shape)


def dispersion(L, dtype, mass, cutoff_type=CUTOFF_NONE, cutoff_sq=np.pi ** 2, bc=None):
    lx, ly, lz = (L, L, L) if np.isscalar(L) else L
    out = np.empty(lx * ly * lz)
    lib().orc_dispersion(out, lx, ly, lz, _bc(bc), dtype, mass, cutoff_type, cutoff_sq)
    return out


def kinetic_xi(L, dtype, mass, dt, cutoff_sq=np.pi ** 2, bc=None):
    lx, ly, lz = (L, L, L) if np.isscalar(L) else L
    out = np.empty(lx * ly * lz)
    lib().orc_kinetic_xi(out, lx, ly, lz, _bc(bc), dtype, mass, dt, cutoff_sq)
    return out


def kinetic_exp(L, tau, mass):
    out = np.empty(L ** 3)
    lib().orc_kinetic_exp(out, L, L, L, tau, mass)
    return out


POT_FORM_NONE, POT_FORM_HARMONIC = 0, 1
POT_TYPE_LIN, POT_TYPE_EXP = 0, 1


def potential_table(L, form, ptype, k, dt):
    """Potential's table (with its 1/V) for spring constants k = (k1, k2, k3)."""
    out = np.empty(L ** 3)
    rc = lib().orc_potential_table(out, L, L, L, form, ptype, float(k[0]), float(k[1]),
                                   float(k[2]), dt)
    assert rc == 0
    return out


def interaction_table(L, mass, itype, couplings, dt, cutoff_sq=np.pi ** 2, bc=None):
    c = np.ascontiguousarray(couplings, dtype=np.float64)
    out = np.empty(L ** 3)
    rc = lib().orc_interaction_table(out, L, L, L, _bc(bc), mass, itype, c, len(c), dt, cutoff_sq)
    return out, rc


def interaction_field(lat: np.ndarray, f: np.ndarray, L: int) -> np.ndarray:
    lat = np.ascontiguousarray(lat, dtype=np.complex128).reshape(-1).view(np.float64)
    out = np.empty(2 * L ** 3)
    assert lib().orc_interaction_field(out, lat, np.ascontiguousarray(f, np.float64), L, L, L) == 0
    return out.view(np.complex128)


def momenta(L, dtype, mass, fermi_energy, cutoff_sq=np.pi ** 2, bc=None, max_count=4096):
    lx, ly, lz = (L, L, L) if np.isscalar(L) else L
    moms = np.zeros(3 * max_count, dtype=np.int32)
    par = np.zeros(max_count, dtype=np.int32)
    om = np.zeros(max_count)
    n = lib().orc_momenta(lx, ly, lz, _bc(bc), dtype, mass, fermi_energy, cutoff_sq, moms, par,
                          om, max_count)
    if n < 0:
        return n, None, None, None
    return n, moms[: 3 * n].reshape(n, 3), par[:n], om[:n]


def lowest_momenta(L: int, n: int) -> np.ndarray:
    out = np.zeros(3 * n, dtype=np.int32)
    assert lib().orc_lowest_momenta(L, n, out) == 0
    return out.reshape(n, 3)


def ranlxs(level: int, seed: int) -> Ranlxs:
    r = Ranlxs()
    assert lib().orc_ranlxs_init(C.byref(r), level, seed) == 0
    return r


def ranlxs_draws(r: Ranlxs, n: int) -> np.ndarray:
    f = lib().orc_ranlxs_next
    return np.array([f(C.byref(r)) for _ in range(n)], dtype=np.float32)


def philox(ctr, key) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(np.ascontiguousarray(ctr, np.uint32),
                            np.ascontiguousarray(key, np.uint32), out)
    return out


def z2_bits(seed: int, config: int, t: int, vol: int, stream: int = 0) -> np.ndarray:
    w = np.zeros((vol + 31) // 32, dtype=np.uint32)
    lib().orc_z2_bits(seed, config, t, stream, vol, w)
    return w


def z2_tilt_phi(seed: int, config: int, t: int, vol: int, thr: int) -> np.ndarray:
    """The tilted Z2 field: site i is +1 where Philox output i (output i % 4 of the block with
    counter (i / 4, t, config, 2)) is below thr, so P(+1) = thr / 2^32.  The importance-sampled
    field of production/one_body-reweight/main.C:73-76 (a Gaussian shifted by phi0 there) in
    Z2 form; the uniform words are the plain sampler's generator read one word per site."""
    u = z2_bits(seed, config, t, 32 * vol, stream=2)
    return np.where(u < np.uint64(thr), 1, -1).astype(np.int8)


def z2_tilt_log_ratio(phis: np.ndarray, q: float) -> np.ndarray:
    """log r_t = sum over t' < t of log(2^-V / (q^n+ (1 - q)^n-)) for fields phis [T][V]."""
    V = phis.shape[1]
    n_plus = (phis == 1).sum(axis=1).astype(np.float64)
    per = n_plus * np.log(0.5 / q) + (V - n_plus) * np.log(0.5 / (1.0 - q))
    return np.concatenate([[0.0], np.cumsum(per)[:-1]])


def z2_memory_phi(seed: int, config: int, T: int, vol: int, a: int, b: int, r: int,
                  stream: int = 3):
    """Z2 fields with memory, slices 0 .. T-1 of one configuration: every site carries
    M = 1024 m, m(t) = sum_{k >= 1} rho^(k-1) phi(t - k) in fixed point (M(0) = 0, M(t + 1) =
    floor((r M + 2^14) / 2^15) + 1024 phi(t)); at slice t site i is +1 where Philox output i
    (output i % 4 of the block with counter (i / 4, t, config, stream)) is below 2^31 + e,
    e = clip(a + b M, -2^25, 2^25).  No counterpart in the reference: an importance sampler
    for lattice.C:81-86's Z2 field, exact like one_body-reweight/main.C:73-76's shifted one.
    Returns phis [T][vol] (int8) and the cumulative log(p / q) before each slice, [T], from
    numpy's log1p."""
    M = np.zeros(vol, dtype=np.int64)
    phis = np.empty((T, vol), dtype=np.int8)
    inc = np.zeros(T)
    for t in range(T):
        u = z2_bits(seed, config, t, 32 * vol, stream=stream).astype(np.int64)
        e = np.clip(a + b * M, -(1 << 25), 1 << 25)
        phi = np.where(u < (1 << 31) + e, 1, -1)
        inc[t] = -np.log1p(phi * e / 2.0 ** 31).sum()
        M = ((r * M + (1 << 14)) >> 15) + 1024 * phi
        phis[t] = phi
    return phis, np.concatenate([[0.0], np.cumsum(inc)[:-1]])


def z2_memory_log_ratio(phis: np.ndarray, a: int, b: int, r: int) -> np.ndarray:
    """The cumulative log(p / q) of given fields phis [T][vol] under fields with memory (a, b,
    r), as z2_memory_phi computes it for its own draws."""
    T, vol = phis.shape
    M = np.zeros(vol, dtype=np.int64)
    inc = np.zeros(T)
    for t in range(T):
        e = np.clip(a + b * M, -(1 << 25), 1 << 25)
        phi = phis[t].astype(np.int64)
        inc[t] = -np.log1p(phi * e / 2.0 ** 31).sum()
        M = ((r * M + (1 << 14)) >> 15) + 1024 * phi
    return np.concatenate([[0.0], np.cumsum(inc)[:-1]])


def bits_to_phi(words: np.ndarray, vol: int) -> np.ndarray:
    phi = np.zeros(vol, dtype=np.int8)
    lib().orc_bits_to_phi(np.ascontiguousarray(words, np.uint32), vol, phi)
    return phi


def ludcmp(a: np.ndarray):
    """Returns (lu, indx, d) for a square complex matrix (copy)."""
    n = a.shape[0]
    buf = np.ascontiguousarray(a, dtype=np.complex128).reshape(-1).view(np.float64).copy()
    indx = np.zeros(n, dtype=np.int32)
    d = C.c_double(0)
    rc = lib().orc_ludcmp(buf, n, indx, C.byref(d))
    if rc:
        raise ZeroDivisionError("singular matrix")
    return buf.view(np.complex128).reshape(n, n), indx, d.value


def lu_det(lu: np.ndarray, d: float):
    n = lu.shape[0]
    la, arg = C.c_double(0), C.c_double(0)
    ph, det = np.zeros(2), np.zeros(2)
    lib().orc_lu_det(np.ascontiguousarray(lu).reshape(-1).view(np.float64), n, d,
                     C.byref(la), ph, C.byref(arg), det)
    return la.value, complex(ph[0], ph[1]), arg.value, complex(det[0], det[1])


def ref_ludcmp(a: np.ndarray):
    """The reference's own ludcmp.C (oracle/_ref) on a contiguous copy; None if not built."""
    if not os.path.exists(REF_LUDCMP_PATH):
        return None
    R = C.CDLL(REF_LUDCMP_PATH)
    R.ref_ludcmp.argtypes = [_dp, C.c_int, _ip, C.POINTER(C.c_double)]
    R.ref_ludcmp.restype = None
    n = a.shape[0]
    buf = np.ascontiguousarray(a, dtype=np.complex128).reshape(-1).view(np.float64).copy()
    indx = np.zeros(n, dtype=np.int32)
    d = C.c_double(0)
    R.ref_ludcmp(buf, n, indx, C.byref(d))
    return buf.view(np.complex128).reshape(n, n), indx, d.value


def overlap(phi: np.ndarray, psi: np.ndarray) -> np.ndarray:
    ns, vol = phi.shape
    nt = psi.shape[0]
    M = np.zeros(2 * ns * nt)
    lib().orc_overlap(np.ascontiguousarray(phi, np.complex128).reshape(-1).view(np.float64),
                      np.ascontiguousarray(psi, np.complex128).reshape(-1).view(np.float64),
                      M, ns, nt, vol)
    return M.view(np.complex128).reshape(ns, nt)


def dawson(x: float) -> float:
    return lib().orc_dawson(float(x))


def opposite_parity_index(L, bc=None) -> np.ndarray:
    lx, ly, lz = (L, L, L) if np.isscalar(L) else L
    out = np.empty(lx * ly * lz, dtype=np.int32)
    lib().orc_opposite_parity_index(lx, ly, lz, _bc(bc), out)
    return out


def two_body_wavefunc(L, wtype, lam, disp1=0, disp2=0, mass1=1.0, mass2=1.0,
                      cutoff_sq=np.pi ** 2, bc=None) -> np.ndarray:
    lx, ly, lz = (L, L, L) if np.isscalar(L) else L
    out = np.empty(lx * ly * lz)
    lib().orc_two_body_wavefunc(out, lx, ly, lz, _bc(bc), wtype, disp1, disp2, mass1, mass2,
                                lam, cutoff_sq)
    return out


def pair_overlap(psi1: np.ndarray, psi2: np.ndarray, w: np.ndarray, opp: np.ndarray):
    """M[s][t] = sum_i w[i] psi1[s][i] psi2[t][opp[i]] (n1 x n2 complex)."""
    n1, vol = psi1.shape
    n2 = psi2.shape[0]
    M = np.zeros(2 * n1 * n2)
    lib().orc_pair_overlap(np.ascontiguousarray(psi1, np.complex128).reshape(-1).view(np.float64),
                           np.ascontiguousarray(psi2, np.complex128).reshape(-1).view(np.float64),
                           np.ascontiguousarray(w, np.float64),
                           np.ascontiguousarray(opp, np.int32), M, n1, n2, vol)
    return M.view(np.complex128).reshape(n1, n2)


def hermite(n: int, x: float) -> np.ndarray:
    h = np.zeros(n)
    lib().orc_hermite(h, n, float(x))
    return h


def sho_source(coeffs: np.ndarray, L: int, l1: float, l2: float, l3: float) -> np.ndarray:
    """OneBody::Set() for SHO sources: coeffs[n1, n2, n3] (complex, n < min(10, L)) -> the
    k-space source vector (V complex)."""
    c = np.zeros((L, L, L), np.complex128)
    n = min(10, L)
    c[:n, :n, :n] = np.asarray(coeffs)[:n, :n, :n]
    out = np.zeros(2 * L ** 3)
    lib().orc_sho_source(out, c.reshape(-1).view(np.float64), L, l1, l2, l3)
    return out.view(np.complex128)


def gauss_field(rng: Ranlxs, vol: int, shift: float = 0.0) -> np.ndarray:
    """Lattice::Refresh(shift) for FIELD_TYPE_GAUSS: vol polar Box-Muller Gaussians + shift."""
    out = np.zeros(vol)
    lib().orc_gauss_field(C.byref(rng), vol, shift, out)
    return out


def run_config(L, N, T, moms, kin, A=0.4, split=SPLIT_KSPACE, minors=False, smear=None,
               field="philox", seed=7012, config=0, rng=None, phi=None, want_M=False,
               want_psi=False, pair_w=None, pair_opp=None, sources=None, pot=None,
               phi_real=None, sinks=None, norms=None, gather=False):
    """One configuration through T slices.  Returns dict of logabs [T, nm], phase [T, nm]
    (complex), argsum [T, nm], optional M [T, N, N] (the pair matrix with pair_w) and
    psi [N, V].  With pair_w / pair_opp: nm = 2N, SlaterDet3's determinants (2n + 1: leading
    (n+1) minor of the pair matrix; 2n: the same with column n from the one-body overlaps)."""
    moms = np.ascontiguousarray(moms, dtype=np.int32).reshape(-1)
    kin = np.ascontiguousarray(kin, dtype=np.float64)
    nm = 2 * N if pair_w is not None else (N if minors else 1)
    la = np.zeros(T * nm)
    ph = np.zeros(2 * T * nm)
    ag = np.zeros(T * nm)
    M = np.zeros(2 * T * N * N) if want_M else None
    ps = np.zeros(2 * N * L ** 3) if want_psi else None
    d = RunDesc()
    d.L, d.N, d.T, d.split, d.minors = L, N, T, split, int(minors)
    d.moms = moms.ctypes.data
    d.kin = kin.ctypes.data
    d.A = A
    d.gather = int(gather)
    sm = None
    if smear is not None:
        sm = np.ascontiguousarray(smear, dtype=np.float64)
        d.smear = sm.ctypes.data
    keep = []
    if sources is not None:  # N x V complex, replacing the plane waves of moms
        src = np.ascontiguousarray(sources, dtype=np.complex128).reshape(-1).view(np.float64)
        keep.append(src)
        d.sources = src.ctypes.data
    if pot is not None:  # Potential table: slices K, P, I, P, K (trapped production order)
        pt = np.ascontiguousarray(pot, dtype=np.float64)
        d.pot = pt.ctypes.data
    if pair_w is not None:
        pw = np.ascontiguousarray(pair_w, dtype=np.float64)
        po = np.ascontiguousarray(pair_opp, dtype=np.int32)
        keep += [pw, po]
        d.pair_w, d.pair_opp = pw.ctypes.data, po.ctypes.data
    if sinks is not None:  # N x V complex, replacing the sources as measurement vectors
        snk = np.ascontiguousarray(sinks, dtype=np.complex128).reshape(-1).view(np.float64)
        keep.append(snk)
        d.sinks = snk.ctypes.data
    if norms is not None:  # T values: Propagator::Normalize(norms[t]) after slice t
        nr = np.ascontiguousarray(norms, dtype=np.float64)
        keep.append(nr)
        d.norms = nr.ctypes.data
    if field == "real":  # T x V real field values
        fr = np.ascontiguousarray(phi_real, dtype=np.float64).reshape(-1)
        keep.append(fr)
        d.field_kind, d.phi_real = FIELD_REAL, fr.ctypes.data
    elif field == "philox":
        d.field_kind, d.seed, d.config = FIELD_PHILOX, seed, config
    elif field == "ranlxs":
        d.field_kind = FIELD_RANLXS
        d.rng = C.cast(C.pointer(rng), C.c_void_p)
    else:
        d.field_kind = FIELD_ARRAY
        arr = np.ascontiguousarray(phi, dtype=np.int8).reshape(-1)
        keep.append(arr)
        d.phi = arr.ctypes.data
    rc = lib().orc_run_config(C.byref(d), la.ctypes.data, ph.ctypes.data, ag.ctypes.data,
                              None if M is None else M.ctypes.data,
                              None if ps is None else ps.ctypes.data)
    assert rc == 0, rc
    out = {"logabs": la.reshape(T, nm), "phase": ph.view(np.complex128).reshape(T, nm),
           "argsum": ag.reshape(T, nm)}
    if want_M:
        out["M"] = M.view(np.complex128).reshape(T, N, N)
    if want_psi:
        out["psi"] = ps.view(np.complex128).reshape(N, L ** 3)
    return out


def exact_c1(L, T, ka, kb, A):
    C_ = np.zeros(T)
    assert lib().orc_exact_c1(L, T, np.ascontiguousarray(ka, np.float64),
                              np.ascontiguousarray(kb, np.float64), A, C_) == 0
    return C_


def time_threads(L, N, slices, moms, kin, A=0.4, split=SPLIT_KSPACE, nthreads=1,
                 seed=7012, config=0):
    """Wall seconds for `nthreads` independent config streams of `slices` slices each."""
    moms = np.ascontiguousarray(moms, dtype=np.int32).reshape(-1)
    kin = np.ascontiguousarray(kin, dtype=np.float64)
    d = RunDesc()
    d.L, d.N, d.T, d.split, d.minors = L, N, slices, split, 0
    d.moms = moms.ctypes.data
    d.kin = kin.ctypes.data
    d.A = A
    d.field_kind, d.seed, d.config = FIELD_PHILOX, seed, config
    return lib().orc_time_threads(C.byref(d), nthreads, slices)


def time_pool(L, N, configs, slices, moms, kin, A=0.4, split=SPLIT_KSPACE, nthreads=1, reps=3,
              seed=7012, config=0, gather=False):
    """The CPU-baseline protocol (orc_time_pool): `nthreads` streams, a one-slice warm-up
    each, then `reps` barrier-synchronised repetitions of `configs` configurations x `slices`
    slices per stream.  Returns the wall seconds of every repetition.  gather: the overlaps
    read at the sink momenta instead of the reference's N^2 full-length products."""
    moms = np.ascontiguousarray(moms, dtype=np.int32).reshape(-1)
    kin = np.ascontiguousarray(kin, dtype=np.float64)
    d = RunDesc()
    d.L, d.N, d.T, d.split, d.minors = L, N, slices, split, 0
    d.gather = int(gather)
    d.moms = moms.ctypes.data
    d.kin = kin.ctypes.data
    d.A = A
    d.field_kind, d.seed, d.config = FIELD_PHILOX, seed, config
    secs = np.zeros(reps)
    rc = lib().orc_time_pool(C.byref(d), nthreads, configs, slices, reps, secs)
    if rc != 0:
        raise RuntimeError("orc_time_pool failed")
    return secs
