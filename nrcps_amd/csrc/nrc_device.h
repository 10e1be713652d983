// nrc_device.h -- device-side building blocks shared by the nrCPS HIP kernels (gfx950).
//
//   * C2<T>            interleaved complex value (re, im) of float or double;
//   * fft_reg<T,L,S>   fully unrolled in-register Stockham DFT of length L, sign S;
//   * twiddles         e^{2 pi i e / L} evaluated at compile time (kTw<L>), folded into the
//                      unrolled transforms as literals.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nrc {

template <typename T>
struct alignas(2 * sizeof(T)) C2 {
  T x, y;
};

template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ C2<T> cscale(C2<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by S*i
template <typename T, int S>
__device__ __forceinline__ C2<T> cmul_i(C2<T> a) {
  return S > 0 ? C2<T>{-a.y, a.x} : C2<T>{a.y, -a.x};
}

// ------------------------------------------------------------------ supported lengths
// Radix plans: 2 = {2}, 4 = {4}, 6 = {3,2}, 8 = {8}, 12 = {4,3}, 16 = {4,4}, 24 = {8,3},
// 32 = {8,4}, 48 = {4,4,3}, 64 = {8,8}.
__host__ __device__ constexpr int fft_nstages(int L) {
  return L == 2 ? 1 : L == 4 ? 1 : L == 6 ? 2 : L == 8 ? 1 : L == 12 ? 2 : L == 16 ? 2
       : L == 24 ? 2 : L == 32 ? 2 : L == 48 ? 3 : L == 64 ? 2 : 0;
}
__host__ __device__ constexpr int fft_radix(int L, int s) {
  return L == 2 ? 2
       : L == 4 ? 4
       : L == 6 ? (s == 0 ? 3 : 2)
       : L == 8 ? 8
       : L == 12 ? (s == 0 ? 4 : 3)
       : L == 16 ? 4
       : L == 24 ? (s == 0 ? 8 : 3)
       : L == 32 ? (s == 0 ? 8 : 4)
       : L == 48 ? (s < 2 ? 4 : 3)
       : L == 64 ? 8 : 0;
}
__host__ __device__ constexpr int fft_ns_before(int L, int s) {
  int n = 1;
  for (int i = 0; i < s; ++i) n *= fft_radix(L, i);
  return n;
}

// ------------------------------------------------------------------ compile-time twiddles
// cos / sin of 2 pi e / L evaluated at compile time: exact integer reduction to the first
// octant, then Taylor series (|error| < 2e-16).  Inside the unrolled transforms e is a
// constant, so every twiddle folds into instruction literals: no tables, no registers held.
constexpr double CX_PI_2 = 1.5707963267948966192313216916398;
__host__ __device__ constexpr double cx_sin_poly(double x) {
  double x2 = x * x, t = x, s = x;
  for (int k = 1; k < 16; ++k) { t *= -x2 / ((2.0 * k) * (2.0 * k + 1.0)); s += t; }
  return s;
}
__host__ __device__ constexpr double cx_cos_poly(double x) {
  double x2 = x * x, t = 1.0, s = 1.0;
  for (int k = 1; k < 16; ++k) { t *= -x2 / ((2.0 * k - 1.0) * (2.0 * k)); s += t; }
  return s;
}
struct CxCS { double c, s; };
__host__ __device__ constexpr CxCS cx_sincos(int e, int L) {
  e = ((e % L) + L) % L;
  const int n = 4 * e, q = n / L, r = n - q * L;  // angle = (pi/2) (q + r/L)
  double cb = 0.0, sb = 0.0;
  if (2 * r <= L) {
    const double th = CX_PI_2 * r / L;
    cb = cx_cos_poly(th);
    sb = cx_sin_poly(th);
  } else {
    const double th = CX_PI_2 * (L - r) / L;
    cb = cx_sin_poly(th);
    sb = cx_cos_poly(th);
  }
  switch (q & 3) {
    case 0: return {cb, sb};
    case 1: return {-sb, cb};
    case 2: return {-cb, -sb};
    default: return {sb, -cb};
  }
}
template <int L>
struct TwTable {
  CxCS w[L > 0 ? L : 1];
};
template <int L>
__host__ __device__ constexpr TwTable<L> make_tw() {
  TwTable<L> t{};
  for (int e = 0; e < L; ++e) t.w[e] = cx_sincos(e, L);
  return t;
}
template <int L>
inline constexpr TwTable<L> kTw = make_tw<L>();

// x * (hi + lo) for an FP32 constant split into its nearest float and remainder.  The small
// product comes first and the large one last, inside one fma, so the remainder reaches the
// result before its single final rounding (added after that rounding it would fall below
// half an ulp and be lost).
template <typename T>
__device__ __forceinline__ T mulk(T x, T hi, T lo) {
  if constexpr (sizeof(T) == 4) return fmaf(x, hi, x * lo);
  else return x * hi;
}

// a * e^{S 2 pi i e / L}; e is a compile-time constant once the caller is unrolled.
// In FP32 the twiddle is applied as a*(hi + lo) with lo the FP32 remainder of the FP64 value,
// the remainder terms summed first and the hi products last (see mulk).  Twiddles rounded to
// FP32 are biased in magnitude: every forward/backward pair would then shrink the orbitals by
// ~4e-8 (L = 32), coherently slice after slice; with the remainders folded in before the
// final rounding the pair is unbiased to ~1e-11.
template <typename T, int L, int S, int E>
__device__ __forceinline__ C2<T> twiddle_c(C2<T> a) {
  if constexpr (E == 0) {
    return a;
  } else if constexpr (2 * E == L) {
    return {-a.x, -a.y};
  } else if constexpr (4 * E == L) {
    return cmul_i<T, S>(a);
  } else if constexpr (4 * E == 3 * L) {
    return cmul_i<T, -S>(a);
  } else {
    constexpr double wc = kTw<L>.w[E].c, ws = (S > 0 ? 1.0 : -1.0) * kTw<L>.w[E].s;
    if constexpr (sizeof(T) == 4) {
      constexpr float hx = (float)wc, hy = (float)ws;
      constexpr float lx = (float)(wc - (double)hx), ly = (float)(ws - (double)hy);
      const float re = fmaf(a.x, hx, fmaf(-a.y, hy, fmaf(a.x, lx, -a.y * ly)));
      const float im = fmaf(a.x, hy, fmaf(a.y, hx, fmaf(a.x, ly, a.y * lx)));
      return {re, im};
    } else {
      return cmul(a, C2<double>{wc, ws});
    }
  }
}

// ------------------------------------------------------------------ small DFTs, w = e^{S 2pi i/R}
template <typename T, int R, int S>
struct Dft;

template <typename T, int S>
struct Dft<T, 2, S> {
  __device__ __forceinline__ static void run(C2<T>* v) {
    C2<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T, int S>
struct Dft<T, 3, S> {
  __device__ __forceinline__ static void run(C2<T>* v) {
    // sin(2 pi / 3) as nearest value plus FP32 remainder
    constexpr double H = 0.8660254037844386467637231707529362;
    const T hh = T(H), hl = T(H - (double)(float)H);
    C2<T> s = cadd(v[1], v[2]);
    C2<T> d = csub(v[1], v[2]);
    C2<T> m = {v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y};
    C2<T> r = cmul_i<T, S>(C2<T>{mulk(d.x, hh, hl), mulk(d.y, hh, hl)});
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

template <typename T, int S>
struct Dft<T, 4, S> {
  __device__ __forceinline__ static void run(C2<T>* v) {
    C2<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    C2<T> s13 = cadd(v[1], v[3]), d13 = cmul_i<T, S>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[1] = cadd(d02, d13);
    v[2] = csub(s02, s13);
    v[3] = csub(d02, d13);
  }
};

template <typename T, int S>
struct Dft<T, 8, S> {
  __device__ __forceinline__ static void run(C2<T>* v) {
    // 1/sqrt(2) as nearest value plus FP32 remainder
    constexpr double R2 = 0.7071067811865475244008443621048490;
    const T rh = T(R2), rl = T(R2 - (double)(float)R2);
    C2<T> e[4] = {v[0], v[2], v[4], v[6]};
    C2<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<T, 4, S>::run(e);
    Dft<T, 4, S>::run(o);
    // o[m] *= w8^m, w8 = e^{S i pi/4}
    C2<T> o1 = {mulk(o[1].x - T(S) * o[1].y, rh, rl), mulk(o[1].y + T(S) * o[1].x, rh, rl)};
    C2<T> o2 = cmul_i<T, S>(o[2]);
    C2<T> o3 = {mulk(-o[3].x - T(S) * o[3].y, rh, rl), mulk(-o[3].y + T(S) * o[3].x, rh, rl)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

// ------------------------------------------------------------------ Stockham stage in registers
// The loops over butterflies (J) and their inputs (RR) are template recursions, so every
// twiddle exponent is a constant expression and its value a literal.
template <typename T, int L, int S, int NR, int K, int TSTEP, int J, int RR, int R>
__device__ __forceinline__ void stockham_load(const C2<T>* in, C2<T>* v) {
  constexpr int E = (RR * K * TSTEP) % L;
  v[RR] = twiddle_c<T, L, S, E>(in[J + RR * NR]);
  if constexpr (RR + 1 < R) stockham_load<T, L, S, NR, K, TSTEP, J, RR + 1, R>(in, v);
}

template <typename T, int L, int S, int STAGE, int J>
__device__ __forceinline__ void stockham_butterfly(const C2<T>* in, C2<T>* out) {
  constexpr int R = fft_radix(L, STAGE);
  constexpr int NS = fft_ns_before(L, STAGE);
  constexpr int NR = L / R;
  constexpr int TSTEP = L / (NS * R);
  constexpr int K = J % NS;
  C2<T> v[R];
  stockham_load<T, L, S, NR, K, TSTEP, J, 0, R>(in, v);
  Dft<T, R, S>::run(v);
  constexpr int D = (J / NS) * NS * R + K;
#pragma unroll
  for (int r = 0; r < R; ++r) out[D + r * NS] = v[r];
  if constexpr (J + 1 < NR) stockham_butterfly<T, L, S, STAGE, J + 1>(in, out);
}

template <typename T, int L, int S, int STAGE>
__device__ __forceinline__ void stockham_stage(const C2<T>* in, C2<T>* out) {
  stockham_butterfly<T, L, S, STAGE, 0>(in, out);
}

// In-place (by value semantics) unnormalised DFT of v[0..L-1]:
// v[k] <- sum_j v[j] e^{S 2 pi i j k / L}.  S = -1 is FFTW_FORWARD.
template <typename T, int L, int S>
__device__ __forceinline__ void fft_reg(C2<T>* v) {
  constexpr int NST = fft_nstages(L);
  static_assert(NST >= 1 && NST <= 3, "unsupported FFT length");
  C2<T> w[L];
  stockham_stage<T, L, S, 0>(v, w);
  if constexpr (NST == 1) {
#pragma unroll
    for (int i = 0; i < L; ++i) v[i] = w[i];
  } else if constexpr (NST == 2) {
    stockham_stage<T, L, S, 1>(w, v);
  } else {
    stockham_stage<T, L, S, 1>(w, v);
    stockham_stage<T, L, S, 2>(v, w);
#pragma unroll
    for (int i = 0; i < L; ++i) v[i] = w[i];
  }
}

// v[i] for a runtime index i, as sum_k [k == i] v[k].  A compare-and-select chain is turned
// back into a dynamically indexed load by the optimiser, which forces the whole array into
// scratch memory; the weighted sum keeps v in registers and is exact for finite values.
template <typename T, int L>
__device__ __forceinline__ C2<T> reg_select(const C2<T>* v, int i) {
  C2<T> r = {T(0), T(0)};
#pragma unroll
  for (int k = 0; k < L; ++k) {
    const T w = (k == i) ? T(1) : T(0);
    r.x += w * v[k].x;
    r.y += w * v[k].y;
  }
  return r;
}

// ------------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11): ten
// rounds of the 4x32 Philox bijection on the counter c under the key (k0, k1).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}

// ------------------------------------------------------------------ proposal threshold
// q(+1) = T / 2^32 of a proposal logit l: T = ceil(s 2^32) clamped to [1, 2^32 - 1], where
// s = 1 / (1 + e) and e = exp(-clamp(l, -30, 30)) by Cody and Waite's reduction
// (n = rint(x log2(e)), r = x - n ln2_hi - n ln2_lo) and the degree-6 Taylor polynomial of
// e^r (|r| <= ln2 / 2, relative error below 2e-7).  Every step is one correctly rounded FP32
// operation with no contraction into fused multiply-adds, so T is reproducible bit for bit
// on any IEEE machine (oracle/convnet.py restates it in numpy float32).
__device__ inline uint32_t proposal_threshold(float l) {
#pragma clang fp contract(off)
  const float x = -fminf(fmaxf(l, -30.f), 30.f);
  const float n = rintf(x * 1.44269504f);
  const float r = (x - n * 0.693145751953125f) - n * 1.42860677e-6f;
  float p = 1.38888889e-3f;
  p = p * r + 8.33333333e-3f;
  p = p * r + 4.16666667e-2f;
  p = p * r + 0.166666667f;
  p = p * r + 0.5f;
  p = p * r + 1.f;
  p = p * r + 1.f;
  const float e = ldexpf(p, (int)n);
  const float sg = 1.f / (1.f + e);
  const float T = ceilf(sg * 4294967296.f);
  return T < 1.f ? 1u : (T >= 4294967296.f ? 4294967295u : (uint32_t)T);
}

// Sum over sites of log(q_uniform / q) = log(2^31 / f), f = T or 2^32 - T, as a running FP64
// product of the f in [1/2, 1) mantissa / exponent form: one log per accumulator at the end.
struct RatioAcc {
  double m = 1.0;
  int e = 0;  // sum of exponents minus 31 per site
  __device__ __forceinline__ void add(uint32_t f) {
    int k;
    m = frexp(m * (double)f, &k);
    e += k - 31;
  }
  __device__ __forceinline__ double value() const {
    return -((double)e * 0.69314718055994531 + log(m));
  }
};

}  // namespace nrc
