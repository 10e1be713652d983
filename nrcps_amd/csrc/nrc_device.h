// nrc_device.h -- device-side building blocks shared by the nrCPS HIP kernels (gfx950).
//
//   * C2<T>            interleaved complex value (re, im) of float or double;
//   * fft_reg<T,L,S>   fully unrolled in-register Stockham DFT of length L, sign S;
//   * twiddle tables   e^{2 pi i e / L} per supported L in __constant__ memory (host fills
//                      them in FP64 with libm at context creation; the FP32 table holds the
//                      correctly rounded FP64 values).
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
// Radix plans: 4 = {4}, 6 = {3,2}, 8 = {8}, 12 = {4,3}, 16 = {4,4}, 24 = {8,3}, 32 = {8,4},
// 48 = {4,4,3}, 64 = {8,8}.
__host__ __device__ constexpr int fft_nstages(int L) {
  return L == 4 ? 1 : L == 6 ? 2 : L == 8 ? 1 : L == 12 ? 2 : L == 16 ? 2 : L == 24 ? 2
       : L == 32 ? 2 : L == 48 ? 3 : L == 64 ? 2 : 0;
}
__host__ __device__ constexpr int fft_radix(int L, int s) {
  return L == 4 ? 4
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
// offsets of each length's table inside the twiddle arrays
__host__ __device__ constexpr int tw_offset(int L) {
  return L == 4 ? 0 : L == 6 ? 4 : L == 8 ? 10 : L == 12 ? 18 : L == 16 ? 30 : L == 24 ? 46
       : L == 32 ? 70 : L == 48 ? 102 : L == 64 ? 150 : -1;
}
constexpr int TW_TOTAL = 214;  // 4+6+8+12+16+24+32+48+64

}  // namespace nrc

// cos, sin of 2 pi e / L for e < L, per supported L at tw_offset(L).  The FP32 table holds the
// nearest floats and c_tw32_lo the floats nearest to the remainders: an FP32 twiddle is
// applied as a*(hi + lo), so |w| = 1 to ~2^-48 and the forward/backward pair of every slice
// does not shrink or grow the orbitals systematically.
extern __constant__ nrc::C2<double> c_tw64[nrc::TW_TOTAL];
extern __constant__ nrc::C2<float> c_tw32[nrc::TW_TOTAL];
extern __constant__ nrc::C2<float> c_tw32_lo[nrc::TW_TOTAL];

namespace nrc {

// x * (hi + lo) for an FP32 constant split into its nearest float and remainder
template <typename T>
__device__ __forceinline__ T mulk(T x, T hi, T lo) {
  if constexpr (sizeof(T) == 4) return fmaf(x, lo, x * hi);
  else return x * hi;
}

// a * e^{S 2 pi i e / L}; e is a compile-time constant once the caller is unrolled
template <typename T, int L, int S>
__device__ __forceinline__ C2<T> twiddle(C2<T> a, int e) {
  if (e == 0) return a;
  if (2 * e == L) return {-a.x, -a.y};
  if (4 * e == L) return cmul_i<T, S>(a);
  if (4 * e == 3 * L) return cmul_i<T, -S>(a);
  if constexpr (sizeof(T) == 4) {
    C2<float> h = c_tw32[tw_offset(L) + e], l = c_tw32_lo[tw_offset(L) + e];
    if (S < 0) { h.y = -h.y; l.y = -l.y; }
    const float re = fmaf(a.x, l.x, fmaf(-a.y, l.y, fmaf(a.x, h.x, -a.y * h.y)));
    const float im = fmaf(a.x, l.y, fmaf(a.y, l.x, fmaf(a.x, h.y, a.y * h.x)));
    return {re, im};
  } else {
    C2<double> w = c_tw64[tw_offset(L) + e];
    w.y = S > 0 ? w.y : -w.y;
    return cmul(a, w);
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
template <typename T, int L, int S, int STAGE>
__device__ __forceinline__ void stockham_stage(const C2<T>* in, C2<T>* out) {
  constexpr int R = fft_radix(L, STAGE);
  constexpr int NS = fft_ns_before(L, STAGE);
  constexpr int NR = L / R;
  constexpr int TSTEP = L / (NS * R);
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const int k = j % NS;
    C2<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = twiddle<T, L, S>(in[j + r * NR], (r * k * TSTEP) % L);
    Dft<T, R, S>::run(v);
    const int d = (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out[d + r * NS] = v[r];
  }
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

}  // namespace nrc
