// fft_axis.hip -- 3D DFTs of any X x Y x Z lattice with every side in {2, 4, 6, 8, 12, 16, 24,
// 32, 48, 64}: the Fourier class for the shapes the fused passes (cubic L <= 32) do not cover,
// non-cubic lattices (src/fourier/noarch/fourier.C:29-39 plans X x Y x Z) and L = 48, 64
// (tests/hamiltonian_performance runs L = 64).
//
// One launch per axis.  A volume is [X][Y][Z], z fastest; along an axis of length N and
// stride s a pencil p = q s + i (q = volume-and-outer index, i < s) starts at q N s + i, so
// the pencils of the whole batch are enumerated without knowing the volume.  Axes with s > 1
// are read by one thread per pencil, lanes on consecutive i (coalesced rows).  The z axis
// (s = 1) is staged through LDS so that loads and stores stay coalesced.
//   * N <= 32: the pencil's DFT runs in registers (fft_reg, as in the fused passes);
//   * N = 48, 64: N = A B with A = 8, and B threads per pencil: each thread takes x[B n1 +
//     n2] (n1 < A) for its n2, runs the A-point DFT, multiplies by w_N^(n2 k1) and parks the
//     result in LDS; then each of A threads takes one k1 and runs the B-point DFT over n2,
//     giving X[k1 + A k2] (Cooley-Tukey with one LDS exchange).
// Every element is read and written once per axis: three HBM round trips per 3D DFT.
#include "nrc_ops.h"

namespace nrc {

namespace {

// runtime twiddle w_N^(S e) from the compile-time table (e < N)
template <int N>
__constant__ TwTable<N> kTwRt = make_tw<N>();

template <typename T, int N, int S>
__device__ __forceinline__ C2<T> twiddle_rt(C2<T> a, int e) {
  const double wc = kTwRt<N>.w[e].c, ws = (S > 0 ? 1.0 : -1.0) * kTwRt<N>.w[e].s;
  if constexpr (sizeof(T) == 4) {
    // FP32: the twiddle as nearest float plus remainder, remainder products first
    const float hx = (float)wc, hy = (float)ws;
    const float lx = (float)(wc - (double)hx), ly = (float)(ws - (double)hy);
    const float re = fmaf(a.x, hx, fmaf(-a.y, hy, fmaf(a.x, lx, -a.y * ly)));
    const float im = fmaf(a.x, hy, fmaf(a.y, hx, fmaf(a.x, ly, a.y * lx)));
    return {re, im};
  } else {
    return cmul(a, C2<double>{wc, ws});
  }
}

constexpr int kRegPencils = 64;  // pencils per workgroup, register path

// N <= 32: one thread per pencil
template <typename T, int N, int S>
__global__ __launch_bounds__(kRegPencils) void k_axis_reg(C2<T>* buf, long npencils, long s) {
  using C = C2<T>;
  __shared__ C sm[kRegPencils * (N + 1)];
  const long p0 = (long)blockIdx.x * kRegPencils;
  const int tid = threadIdx.x;
  const long p = p0 + tid;
  C v[N];
  if (s == 1) {  // the tile's pencils are contiguous: N * (pencils in the tile) elements
    const long left = npencils - p0;
    const int np = (int)(left < kRegPencils ? left : kRegPencils);
    C* g = buf + p0 * N;
    for (int e = tid; e < np * N; e += kRegPencils) sm[(e / N) * (N + 1) + e % N] = g[e];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = sm[tid * (N + 1) + j];
    fft_reg<T, N, S>(v);
#pragma unroll
    for (int j = 0; j < N; ++j) sm[tid * (N + 1) + j] = v[j];
    __syncthreads();
    for (int e = tid; e < np * N; e += kRegPencils) g[e] = sm[(e / N) * (N + 1) + e % N];
    return;
  }
  if (p >= npencils) return;
  const long q = p / s, i = p - q * s;
  C* g = buf + q * N * s + i;
#pragma unroll
  for (int j = 0; j < N; ++j) v[j] = g[(long)j * s];
  fft_reg<T, N, S>(v);
#pragma unroll
  for (int j = 0; j < N; ++j) g[(long)j * s] = v[j];
}

// N = A * B (A = 8; B = 6 or 8): B threads per pencil in step 1, A in step 2
template <typename T, int N, int S>
__global__ __launch_bounds__(256) void k_axis_two(C2<T>* buf, long npencils, long s) {
  using C = C2<T>;
  constexpr int A = 8, B = N / A, TPP = A > B ? A : B, P = 256 / TPP, LD = N + 1;
  static_assert(A * B == N, "N must be a multiple of 8");
  __shared__ C sm[P * LD];
  const long p0 = (long)blockIdx.x * P;
  const int tid = threadIdx.x;
  const long left = npencils - p0;
  const int np = (int)(left < P ? left : P);
  // load the tile: z axis contiguous, strided axes with consecutive pencils on consecutive
  // threads
  if (s == 1) {
    const C* g = buf + p0 * N;
    for (int e = tid; e < np * N; e += 256) sm[(e / N) * LD + e % N] = g[e];
  } else {
    for (int e = tid; e < np * N; e += 256) {
      const int pl = e % np, j = e / np;
      const long p = p0 + pl, q = p / s, i = p - q * s;
      sm[pl * LD + j] = buf[q * N * s + i + (long)j * s];
    }
  }
  __syncthreads();
  const int pl = tid / TPP, r = tid - pl * TPP;
  C* t = sm + pl * LD;
  const bool live = pl < np;
  // step 1: n2 = r < B; A-point DFT over n1 of x[B n1 + n2], then w_N^(n2 k1)
  if (live && r < B) {
    C v[A];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) v[n1] = t[B * n1 + r];
    fft_reg<T, A, S>(v);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) t[B * k1 + r] = twiddle_rt<T, N, S>(v[k1], (r * k1) % N);
  }
  __syncthreads();
  // step 2: k1 = r < A; B-point DFT over n2 of Y[n2][k1] (slot B k1 + n2) -> X[k1 + A k2]
  C v[B];
  if (live && r < A) {
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) v[n2] = t[B * r + n2];
    fft_reg<T, B, S>(v);
  }
  __syncthreads();
  if (live && r < A) {
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) t[r + A * k2] = v[k2];
  }
  __syncthreads();
  if (s == 1) {
    C* g = buf + p0 * N;
    for (int e = tid; e < np * N; e += 256) g[e] = sm[(e / N) * LD + e % N];
  } else {
    for (int e = tid; e < np * N; e += 256) {
      const int q2 = e % np, j = e / np;
      const long p = p0 + q2, q = p / s, i = p - q * s;
      buf[q * N * s + i + (long)j * s] = sm[q2 * LD + j];
    }
  }
}

template <typename T, int N, int S>
hipError_t launch_axis_n(void* buf, long npencils, long s, hipStream_t st) {
  if constexpr (N <= 32) {
    const unsigned blocks = (unsigned)((npencils + kRegPencils - 1) / kRegPencils);
    hipLaunchKernelGGL((k_axis_reg<T, N, S>), dim3(blocks), dim3(kRegPencils), 0, st,
                       (C2<T>*)buf, npencils, s);
  } else {
    constexpr int P = 256 / (N / 8 > 8 ? N / 8 : 8);
    const unsigned blocks = (unsigned)((npencils + P - 1) / P);
    hipLaunchKernelGGL((k_axis_two<T, N, S>), dim3(blocks), dim3(256), 0, st, (C2<T>*)buf,
                       npencils, s);
  }
  return hipGetLastError();
}

template <typename T, int S>
hipError_t launch_axis_t(void* buf, int n, long npencils, long s, hipStream_t st) {
  switch (n) {
    case 2: return launch_axis_n<T, 2, S>(buf, npencils, s, st);
    case 4: return launch_axis_n<T, 4, S>(buf, npencils, s, st);
    case 6: return launch_axis_n<T, 6, S>(buf, npencils, s, st);
    case 8: return launch_axis_n<T, 8, S>(buf, npencils, s, st);
    case 12: return launch_axis_n<T, 12, S>(buf, npencils, s, st);
    case 16: return launch_axis_n<T, 16, S>(buf, npencils, s, st);
    case 24: return launch_axis_n<T, 24, S>(buf, npencils, s, st);
    case 32: return launch_axis_n<T, 32, S>(buf, npencils, s, st);
    case 48: return launch_axis_n<T, 48, S>(buf, npencils, s, st);
    case 64: return launch_axis_n<T, 64, S>(buf, npencils, s, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

bool axis_supported(int n) {
  return n == 2 || n == 4 || n == 6 || n == 8 || n == 12 || n == 16 || n == 24 || n == 32 ||
         n == 48 || n == 64;
}

// unnormalised 3D DFT of `batch` volumes [lx][ly][lz] in place; sign -1 is FFTW_FORWARD
hipError_t launch_fft3d_axes(int prec, void* buf, int lx, int ly, int lz, long batch, int sign,
                             hipStream_t st) {
  if (batch <= 0) return hipSuccess;
  const long V = (long)lx * ly * lz;
  const int len[3] = {lz, ly, lx};
  const long stride[3] = {1, lz, (long)ly * lz};
  for (int a = 0; a < 3; ++a) {
    const long np = batch * (V / len[a]);
    hipError_t e;
    if (prec == NRC_PREC_F64)
      e = sign < 0 ? launch_axis_t<double, -1>(buf, len[a], np, stride[a], st)
                   : launch_axis_t<double, 1>(buf, len[a], np, stride[a], st);
    else
      e = sign < 0 ? launch_axis_t<float, -1>(buf, len[a], np, stride[a], st)
                   : launch_axis_t<float, 1>(buf, len[a], np, stride[a], st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace nrc
