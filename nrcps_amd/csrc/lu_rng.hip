// lu_rng.hip -- batched FP64 LU log-determinants, dense overlaps, Philox Z2 fields.
#include "nrc_device.h"
#include "nrc_internal.h"

namespace nrc {

// ------------------------------------------------------------------ Philox Z2 fields
// Word w of configuration c's field at slice t is output (w % 4) of the Philox4x32-10 block
// with counter (w / 4, t, c, stream) and key (seed lo, seed hi); bit j of word w is site
// 32 w + j, phi = +1 where set.
// par != null: seed = par[0] and config0 = par[1] are read from device memory, so that a
// captured graph of these launches serves any seed and configuration range
__global__ void k_z2_bits(uint64_t seed, uint32_t config0, int nconf, uint32_t t,
                          uint32_t stream, int nw, uint32_t* words, const uint64_t* par) {
  if (par) {
    seed = par[0];
    config0 = (uint32_t)par[1];
  }
  const int ng = (nw + 3) / 4;
  const long n = (long)nconf * ng;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i / ng), g = (int)(i - (long)c * ng);
    uint32_t ctr[4] = {(uint32_t)g, t, config0 + (uint32_t)c, stream};
    philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t* out = words + (long)c * nw + 4 * g;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (4 * g + q < nw) out[q] = ctr[q];
  }
}

hipError_t launch_z2_bits(uint64_t seed, uint32_t config0, int nconf, uint32_t t,
                          uint32_t stream, int vol, uint32_t* words, hipStream_t s,
                          const uint64_t* par) {
  const int nw = (vol + 31) / 32;
  const long n = (long)nconf * ((nw + 3) / 4);
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_z2_bits, dim3((unsigned)blocks), dim3(256), 0, s, seed, config0, nconf,
                     t, stream, nw, words, par);
  return hipGetLastError();
}

// Tilted Z2 fields: site i of configuration c at slice t is +1 where the Philox output
// u_i < thr, u_i being output (i % 4) of the block with counter (i / 4, t, c, stream): P(+1) =
// thr / 2^32 exactly.  One workgroup per configuration; each thread builds whole words (eight
// Philox blocks per word) and fsum[c] receives the field sum 2 n(+1) - vol, from which the
// sampling ratio of the slice follows.
__global__ __launch_bounds__(256) void k_z2_tilt(uint64_t seed, uint32_t config0, int nconf,
                                                 uint32_t t, uint32_t stream, int vol,
                                                 uint32_t thr, uint32_t* words, int* fsum,
                                                 const uint64_t* par) {
  __shared__ int red[4];
  if (par) {
    seed = par[0];
    config0 = (uint32_t)par[1];
  }
  const int c = blockIdx.x;
  const int nw = (vol + 31) / 32;
  int pop = 0;
  for (int w = threadIdx.x; w < nw; w += 256) {
    uint32_t bits = 0;
#pragma unroll 2
    for (int b = 0; b < 8; ++b) {
      uint32_t ctr[4] = {(uint32_t)(8 * w + b), t, config0 + (uint32_t)c, stream};
      philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
      for (int q = 0; q < 4; ++q) bits |= (uint32_t)(ctr[q] < thr) << (4 * b + q);
    }
    words[(long)c * nw + w] = bits;
    const int valid = vol - 32 * w;
    pop += __popc(valid >= 32 ? bits : bits & ((1u << valid) - 1u));
  }
  for (int off = 32; off >= 1; off >>= 1) pop += __shfl_xor(pop, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = pop;
  __syncthreads();
  if (threadIdx.x == 0) fsum[c] = 2 * (red[0] + red[1] + red[2] + red[3]) - vol;
}

hipError_t launch_z2_tilt(uint64_t seed, uint32_t config0, int nconf, uint32_t t,
                          uint32_t stream, int vol, uint32_t thr, uint32_t* words, int* fsum,
                          hipStream_t s, const uint64_t* par) {
  if (nconf < 1) return hipSuccess;
  hipLaunchKernelGGL(k_z2_tilt, dim3((unsigned)nconf), dim3(256), 0, s, seed, config0, nconf,
                     t, stream, vol, thr, words, fsum, par);
  return hipGetLastError();
}

// Z2 fields with memory.  Every site carries a memory m, the decayed sum of its own past
// field, m(t) = sum_{k >= 1} rho^(k-1) phi(t - k), held in fixed point M = 1024 m (int16):
//   M(0) = 0,   M(t + 1) = floor((r M(t) + 2^14) / 2^15) + 1024 phi(t),   rho = r / 2^15.
// At slice t the site is +1 with probability q = thr / 2^32, thr = 2^31 + e,
// e = clamp(a + b M(t), -2^25, 2^25): q = (1 + eps) / 2 with eps = e / 2^31 = lambda + theta m
// (lambda = a / 2^31, theta = b / 2^21) within |eps| <= 2^-6.  All of it is integer
// arithmetic, so the draw is reproducible bit for bit, and the slice's sampling ratio
// log(p / q) = sum over sites of log((1/2) / q(phi)) = -sum log1p(phi eps) is exact up to
// its FP64 evaluation: log1p by its Taylor series to x^10 (|x| <= 2^-6, truncation below
// 2^-60 |x|).
constexpr int kMemEMax = 1 << 25;

__device__ __forceinline__ int mem_drive(int a, int b, int M) {
  long long e = (long long)a + (long long)b * M;
  e = e > kMemEMax ? kMemEMax : (e < -kMemEMax ? -kMemEMax : e);
  return (int)e;
}

__device__ __forceinline__ int mem_next(int r, int M, bool up) {
  return ((r * M + (1 << 14)) >> 15) + (up ? 1024 : -1024);
}

// -log1p(x) for |x| <= 2^-6: x (-1 + x (1/2 + x (-1/3 + ...)))
__device__ __forceinline__ double neg_log1p_small(double x) {
  double s = 1.0 / 10.0;
  s = fma(x, s, -1.0 / 9.0);
  s = fma(x, s, 1.0 / 8.0);
  s = fma(x, s, -1.0 / 7.0);
  s = fma(x, s, 1.0 / 6.0);
  s = fma(x, s, -1.0 / 5.0);
  s = fma(x, s, 1.0 / 4.0);
  s = fma(x, s, -1.0 / 3.0);
  s = fma(x, s, 1.0 / 2.0);
  s = fma(x, s, -1.0);
  return x * s;
}

// sum over the workgroup (256 threads) of one FP64 value per thread, in a fixed order
__device__ __forceinline__ double block_sum256(double v, double* red) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double s = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return s;
}

// The sampler: slice t of configurations config0 .. config0 + nconf - 1, one workgroup per
// configuration, each thread whole words (eight Philox blocks of counter (i / 4, t, c,
// stream) per word, output i % 4 for site i, as the tilted sampler).  mem[c] holds the
// configuration's M, 32 per word (t = 0: not read, taken as 0); lr[c] receives the slice's
// log(p / q).
__global__ __launch_bounds__(256) void k_z2_memory(uint64_t seed, uint32_t config0, uint32_t t,
                                                   uint32_t stream, int vol, int a, int b,
                                                   int r, uint32_t* words, int16_t* mem,
                                                   double* lr, const uint64_t* par) {
  __shared__ double red[4];
  if (par) {
    seed = par[0];
    config0 = (uint32_t)par[1];
  }
  const int c = blockIdx.x;
  const int nw = (vol + 31) / 32;
  double acc = 0.0;
  for (int w = threadIdx.x; w < nw; w += 256) {
    uint4* m = reinterpret_cast<uint4*>(mem + ((long)c * nw + w) * 32);
    uint32_t mp[16];  // M of sites 2j (low half) and 2j + 1 (high half)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint4 v = t == 0 ? make_uint4(0u, 0u, 0u, 0u) : m[k];
      mp[4 * k] = v.x;
      mp[4 * k + 1] = v.y;
      mp[4 * k + 2] = v.z;
      mp[4 * k + 3] = v.w;
    }
    const int valid = vol - 32 * w;
    uint32_t bits = 0;
#pragma unroll
    for (int blk = 0; blk < 8; ++blk) {
      uint32_t ctr[4] = {(uint32_t)(8 * w + blk), t, config0 + (uint32_t)c, stream};
      philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = 4 * blk + q, sh = 16 * (i & 1);
        const int M = (int16_t)(mp[i >> 1] >> sh);
        const int e = mem_drive(a, b, M);
        const bool up = ctr[q] < (uint32_t)((1u << 31) + (uint32_t)e);
        bits |= (uint32_t)up << i;
        const double x = (up ? 0x1p-31 : -0x1p-31) * (double)e;
        if (i < valid) acc += neg_log1p_small(x);
        const uint32_t nm = (uint32_t)(uint16_t)(int16_t)mem_next(r, M, up);
        mp[i >> 1] = (mp[i >> 1] & (0xffff0000u >> sh)) | (nm << sh);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      m[k] = make_uint4(mp[4 * k], mp[4 * k + 1], mp[4 * k + 2], mp[4 * k + 3]);
    words[(long)c * nw + w] = bits;
  }
  const double s = block_sum256(acc, red);
  if (threadIdx.x == 0) lr[c] = s;
}

hipError_t launch_z2_memory(uint64_t seed, uint32_t config0, int nconf, uint32_t t,
                            uint32_t stream, int vol, int a, int b, int r, uint32_t* words,
                            int16_t* mem, double* lr, hipStream_t s, const uint64_t* par) {
  if (nconf < 1) return hipSuccess;
  hipLaunchKernelGGL(k_z2_memory, dim3((unsigned)nconf), dim3(256), 0, s, seed, config0, t,
                     stream, vol, a, b, r, words, mem, lr, par);
  return hipGetLastError();
}

// The same proposal evaluated on the plain sampler's fields (stream 0, word w of slice t =
// output w % 4 of the block (w / 4, t, c, 0)): out[t][c] = the cumulative log(p / q) of
// slices 0 .. t - 1 (out[0] = 0).  One workgroup per configuration; thread j owns words j,
// j + 256, .. (WPT of them) and walks their memory over t in registers (two sites per
// register).
constexpr int kMemEvalMaxVol = 32768;

template <int WPT>
__global__ __launch_bounds__(256) void k_z2_memory_eval(uint64_t seed, uint32_t config0, int T,
                                                        int vol, int a, int b, int r,
                                                        double* out, int nconf) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  const int nw = (vol + 31) / 32;
  uint32_t mp[WPT][16];
#pragma unroll
  for (int k = 0; k < WPT; ++k)
#pragma unroll
    for (int j = 0; j < 16; ++j) mp[k][j] = 0u;
  double run = 0.0;
  for (int t = 0; t < T; ++t) {
    if (threadIdx.x == 0) out[(long)t * nconf + c] = run;
    if (t == T - 1) break;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
      const int w = threadIdx.x + 256 * k;
      if (w < nw) {
        uint32_t ctr[4] = {(uint32_t)(w >> 2), (uint32_t)t, config0 + (uint32_t)c, 0u};
        philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
        const int q = w & 3;
        const uint32_t word = q == 0 ? ctr[0] : q == 1 ? ctr[1] : q == 2 ? ctr[2] : ctr[3];
        const int valid = vol - 32 * w;
#pragma unroll
        for (int i = 0; i < 32; ++i) {
          const int sh = 16 * (i & 1);
          const int M = (int16_t)(mp[k][i >> 1] >> sh);
          const int e = mem_drive(a, b, M);
          const bool up = (word >> i) & 1u;
          const double x = (up ? 0x1p-31 : -0x1p-31) * (double)e;
          if (i < valid) acc += neg_log1p_small(x);
          const uint32_t nm = (uint32_t)(uint16_t)(int16_t)mem_next(r, M, up);
          mp[k][i >> 1] = (mp[k][i >> 1] & (0xffff0000u >> sh)) | (nm << sh);
        }
      }
    }
    run += block_sum256(acc, red);
  }
}

hipError_t launch_z2_memory_eval(uint64_t seed, uint32_t config0, int nconf, int T, int vol,
                                 int a, int b, int r, double* out, hipStream_t s) {
  if (nconf < 1 || T < 1) return hipSuccess;
  if (vol < 1 || ((vol + 31) / 32) * 32 > kMemEvalMaxVol) return hipErrorInvalidValue;
  const int nw = (vol + 31) / 32;
  if (nw <= 256)
    hipLaunchKernelGGL(k_z2_memory_eval<1>, dim3((unsigned)nconf), dim3(256), 0, s, seed,
                       config0, T, vol, a, b, r, out, nconf);
  else if (nw <= 512)
    hipLaunchKernelGGL(k_z2_memory_eval<2>, dim3((unsigned)nconf), dim3(256), 0, s, seed,
                       config0, T, vol, a, b, r, out, nconf);
  else
    hipLaunchKernelGGL(k_z2_memory_eval<4>, dim3((unsigned)nconf), dim3(256), 0, s, seed,
                       config0, T, vol, a, b, r, out, nconf);
  return hipGetLastError();
}

// ------------------------------------------------------------------ batched LU
// One wavefront per matrix, lane i owns row i, the matrix sits in LDS column-major so that
// lanes touch consecutive addresses.  The factorisation is right-looking: at step k the
// column below the pivot is scaled by the reciprocal pivot and the trailing block gets
// a_ij -= l_ik u_kj.  Every a_ij therefore receives its updates in the order k = 0, 1, ...,
// exactly the order of the Crout dot products in src/utils/ludcmp.C:32-40, and the pivot
// is chosen as there: implicit row scaling vv_i = 1 / max_j |a_ij| and the LAST row with the
// largest vv_i |a_ik| (the reference compares with >=, ludcmp.C:43).  A zero pivot is
// replaced by 1e-20 (ludcmp.C:57).  An all-zero row sets status = 1 (the reference exits).
//
// Every floating-point operation is the one the reference executes, in the same order and
// without fused multiply-adds, so that pivot ties break the same way and the factors are
// bit-identical: |z| is the corrected hypot of C. F. Borges ("An improved algorithm for
// hypot(a,b)", 2019; the non-FMA kernel with the 2^-600 / 2^511 / 2^-511 / 2^-54 scaling
// thresholds of the host C library's hypot, which std::abs(complex<double>) calls), the
// reciprocal pivot is Smith's complex division (the host compiler's complex division), and
// complex products are (ac - bd, ad + bc).
#pragma clang fp contract(off)

__device__ __forceinline__ double hypot_kernel(double ax, double ay) {
  double h = sqrt(ax * ax + ay * ay);
  double t1, t2;
  if (h <= 2.0 * ay) {
    const double delta = h - ay;
    t1 = ax * (2.0 * delta - ax);
    t2 = (delta - 2.0 * (ax - ay)) * delta;
  } else {
    const double delta = h - ax;
    t1 = 2.0 * delta * (ax - 2.0 * ay);
    t2 = (4.0 * delta - ay) * ay + delta * delta;
  }
  h -= (t1 + t2) / (2.0 * h);
  return h;
}

__device__ double ref_abs(double x, double y) {
  if (isinf(x) || isinf(y)) return INFINITY;
  if (isnan(x) || isnan(y)) return x + y;
  x = fabs(x);
  y = fabs(y);
  double ax = x < y ? y : x;
  const double ay = x < y ? x : y;
  const double SCALE = 0x1p-600, LARGE = 0x1p+511, TINY = 0x1p-511, EPS = 0x1p-54;
  if (ax > LARGE) {
    if (ay <= ax * EPS) return ax + ay;
    return hypot_kernel(ax * SCALE, ay * SCALE) / SCALE;
  }
  if (ay < TINY) {
    if (ax >= ay / EPS) return ax + ay;
    ax = hypot_kernel(ax / SCALE, ay / SCALE) * SCALE;
    return ax;
  }
  if (ay <= ax * EPS) return ax + ay;
  return hypot_kernel(ax, ay);
}

__device__ __forceinline__ double cabs2(double2 a) { return ref_abs(a.x, a.y); }

// the maximum over the active lanes of a wavefront (the ROCm device library's DPP reduction)
extern "C" __device__ __attribute__((const)) double __ockl_wfred_max_f64(double);
extern "C" __device__ __attribute__((const)) int __ockl_wfred_max_i32(int);

// lane l's value of x (l uniform)
__device__ __forceinline__ double readlane_d(double x, int l) {
  const unsigned long long b = __double_as_longlong(x);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, l);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// (1 + 0i) / z by Smith's method
__device__ __forceinline__ double2 crecip(double2 z) {
  const double a = 1.0, b = 0.0, c = z.x, d = z.y;
  double ratio, denom, x, y;
  if (fabs(c) < fabs(d)) {
    ratio = c / d;
    denom = (c * ratio) + d;
    x = ((a * ratio) + b) / denom;
    y = ((b * ratio) - a) / denom;
  } else {
    ratio = d / c;
    denom = (d * ratio) + c;
    x = ((b * ratio) + a) / denom;
    y = (b - (a * ratio)) / denom;
  }
  return make_double2(x, y);
}

__device__ __forceinline__ double2 cmuld(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ __launch_bounds__(64) void k_lu_det(const double2* __restrict__ M, int n, int ldm,
                                               int ncfg, int minors, double* logabs,
                                               double2* phase, double* argsum, int* status,
                                               double2* lu_out, int* indx_out, double* d_out,
                                               LuOpts opt) {
  extern __shared__ double2 a[];
  const int nm = minors ? n : 1;
  // minors: the largest of every configuration dispatched first, the smallest last, so that
  // the longest factorizations do not start at the end of the launch
  const long blk = blockIdx.x;
  const int m = nm - 1 - (int)(blk / ncfg);
  const long cfg = blk - (long)(nm - 1 - m) * ncfg;
  const int sz = minors ? m + 1 : n;
  const int i = threadIdx.x;
  const double2* Mc = M + cfg * (long)n * ldm;
  {  // the minor, coalesced: element e = r sz + c, four loads in flight per lane
    const int tot = sz * sz;
    const bool rp = opt.repl != nullptr;
    for (int e0 = 0; e0 < tot; e0 += 256) {
      double2 v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = e0 + 64 * q + i;
        if (e < tot) {
          const int r = e / sz, c = e - r * sz;
          if (!(rp && c == sz - 1)) v[q] = Mc[(long)r * ldm + c];
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = e0 + 64 * q + i;
        if (e < tot) {
          const int r = e / sz, c = e - r * sz;
          if (!(rp && c == sz - 1)) a[c * sz + r] = v[q];
        }
      }
    }
    if (rp && i < sz)
      a[(sz - 1) * sz + i] = opt.repl[cfg * (long)n * opt.ldr + (long)(sz - 1) * opt.ldr + i];
  }
  __syncthreads();
  // Row scale vv_i = 1 / max_j |a_ij| with the exact |.| (ref_abs).  The maximum is found
  // among candidates picked by the squared magnitudes, computed to within 2^-51 relative: an
  // entry whose square is below (1 - 2^-40) of the row's largest square has |z|, and so
  // ref_abs(z) (within an ulp of |z|), below the largest entry's ref_abs.  A row with one
  // candidate takes a single ref_abs; a row with several, or whose largest square lies outside
  // [2^-900, 2^900] (underflow, overflow, infinities, an all-zero row), takes every entry's.
  double vv = 0.0;
  int bad = 0;
  bool slow = false;
  double big = 0.0;
  if (i < sz) {
    double b2 = 0.0, s2 = 0.0;
    int jb = 0;
    for (int j = 0; j < sz; ++j) {
      const double2 z = a[j * sz + i];
      const double m2 = z.x * z.x + z.y * z.y;
      if (m2 > b2) {
        s2 = b2;
        b2 = m2;
        jb = j;
      } else if (m2 > s2) {
        s2 = m2;
      }
    }
    if (!(b2 >= 0x1p-900 && b2 <= 0x1p+900) || s2 >= b2 * (1.0 - 0x1p-40)) slow = true;
    else big = cabs2(a[jb * sz + i]);
  }
  if (__any(slow) && slow) {
    big = 0.0;
    for (int j = 0; j < sz; ++j) {
      const double t = cabs2(a[j * sz + i]);
      if (t > big) big = t;
    }
  }
  if (i < sz) {
    if (big == 0.0) bad = 1;
    else vv = 1.0 / big;
  }
  bad = __any(bad);
  double dsign = 1.0, la = 0.0, as = 0.0;
  double2 ph = make_double2(1.0, 0.0);
  int my_indx = 0;
  // lane k keeps pivot k; the logs, arguments and unit phases of all pivots are taken after
  // the elimination, one pivot per lane, and summed across the wavefront
  double2 my_piv = make_double2(1.0, 0.0);
  double my_mag = 1.0;
  // Rows stay in their lanes; pos is the row's position in the reference's row order (a row
  // swap exchanges two positions), and the rows already pivoted take no further part.
  int pos = i;
  bool done = i >= sz;
  // The pivot search by squared magnitudes: b2 = vv^2 (x^2 + y^2) is within a few ulps of
  // (vv |z|)^2, so a row whose b2 lies below (1 - 2^-40) of the largest has vv ref_abs(z)
  // below that row's, rounding included.  When exactly one row reaches that band, and every
  // square and product lies in [2^-900, 2^900] (no underflow or overflow can reorder them),
  // that row is the reference's pivot; otherwise the step compares vv ref_abs(z) exactly.
  // |pivot| itself is only needed for log|det|, after the elimination.
  const double vv2 = vv * vv;
  const bool fast_rows = !__any(i < sz && !(vv2 >= 0x1p-900 && vv2 <= 0x1p+900));
  if (!bad) {
    for (int k = 0; k < sz; ++k) {
      // the pivot: the largest vv |a_ik| over the rows at positions >= k, the last position
      // among equals (the reference's >=): the wavefront maximum, then among the lanes holding
      // it the highest position
      const double2 z = done ? make_double2(0.0, 0.0) : a[k * sz + i];
      const double m2 = z.x * z.x + z.y * z.y;
      const bool zero = z.x == 0.0 && z.y == 0.0;
      const bool in_range = done || zero || (m2 >= 0x1p-900 && m2 <= 0x1p+900);
      const double b2 = done ? -1.0 : vv2 * m2;
      const double mx2 = __ockl_wfred_max_f64(b2);
      const unsigned long long cand = __ballot(!done && b2 >= mx2 * (1.0 - 0x1p-40));
      int p;
      if (fast_rows && mx2 >= 0x1p-900 && mx2 <= 0x1p+900 && __popcll(cand) == 1 &&
          !__any(!in_range)) {
        p = __ffsll((long long)cand) - 1;
      } else {
        const double h = done ? 0.0 : cabs2(z);
        const double bv = done ? -1.0 : vv * h;
        const double mx = __ockl_wfred_max_f64(bv);
        const bool hit = !done && bv == mx;
        const unsigned long long tie = __ballot(hit);
        p = tie ? __ffsll((long long)tie) - 1 : 0;
        if (__popcll(tie) > 1) {
          const int best = __ockl_wfred_max_i32(hit ? pos : -1);
          p = __ffsll((long long)__ballot(hit && pos == best)) - 1;
        }
      }
      const int imax = __builtin_amdgcn_readlane(pos, p);
      if (i == k) my_indx = imax;
      if (imax != k) {
        if (pos == k) pos = imax;
        dsign = -dsign;
      }
      if (i == p) {
        pos = k;
        done = true;
      }
      double2 piv = make_double2(readlane_d(z.x, p), readlane_d(z.y, p));
      if (piv.x == 0.0 && piv.y == 0.0) {  // |pivot| = 0 exactly
        piv = make_double2(1.0e-20, 0.0);
        if (i == p) a[k * sz + p] = piv;
      }
      if (i == k) my_piv = piv;
      const double2 r = crecip(piv);
      if (!done) {
        const double2 l = cmuld(z, r);
        a[k * sz + i] = l;
        // eight columns at a time, their LDS reads issued together (the pivot row's a
        // broadcast read, row i this lane's own), then the remaining columns one by one
        int j = k + 1;
        for (; j + 8 <= sz; j += 8) {
          double2 u[8], x[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            u[q] = a[(j + q) * sz + p];
            x[q] = a[(j + q) * sz + i];
          }
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const double2 pr = cmuld(l, u[q]);
            x[q].x = x[q].x - pr.x;
            x[q].y = x[q].y - pr.y;
            a[(j + q) * sz + i] = x[q];
          }
        }
        for (; j < sz; ++j) {
          const double2 pr = cmuld(l, a[j * sz + p]);
          double2 x = a[j * sz + i];
          x.x = x.x - pr.x;
          x.y = x.y - pr.y;
          a[j * sz + i] = x;
        }
      }
      __syncthreads();
    }
    if (i < sz) {
      my_mag = cabs2(my_piv);
      la = log(my_mag);
      ph = make_double2(my_piv.x / my_mag, my_piv.y / my_mag);
      as = atan2(my_piv.y, my_piv.x);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      la += __shfl_xor(la, off);
      as += __shfl_xor(as, off);
      ph = cmuld(ph, make_double2(__shfl_xor(ph.x, off), __shfl_xor(ph.y, off)));
    }
  }
  const long o = cfg * (opt.onm ? opt.onm : nm) + (long)m * opt.ostep + opt.oofs;
  if (lu_out && i < sz) {  // row i of the factors is at position pos
    double2* out = lu_out + cfg * (long)n * ldm;
    for (int j = 0; j < sz; ++j) out[(long)pos * ldm + j] = a[j * sz + i];
    indx_out[cfg * (long)n + i] = my_indx;
    if (i == 0) d_out[cfg] = dsign;
  }
  if (i == 0) {
    if (status) status[o] = bad ? 1 : 0;
    const double ncol = opt.lsrc_cols * sz - (opt.repl ? opt.lsrc_cols - 1 : 0);
    const long cs = opt.lsrc_period > 0 ? cfg % opt.lsrc_period : cfg;
    if (logabs) logabs[o] = bad ? -INFINITY : (opt.lsrc ? la - ncol * opt.lsrc[cs] : la);
    if (phase) phase[o] = bad ? make_double2(0.0, 0.0) : make_double2(dsign * ph.x, dsign * ph.y);
    if (argsum) argsum[o] = bad ? 0.0 : as + (dsign < 0.0 ? 3.14159265358979323846 : 0.0);
  }
}

__global__ void k_ref_abs(const double2* z, double* out, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    out[i] = cabs2(z[i]);
}

#pragma clang fp contract(on)

hipError_t launch_ref_abs(const double2* z, double* out, long n, hipStream_t s) {
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_ref_abs, dim3((unsigned)blocks), dim3(256), 0, s, z, out, n);
  return hipGetLastError();
}

hipError_t launch_lu_det(const double2* M, int n, int ldm, int ncfg, int minors,
                         double* logabs, double2* phase, double* argsum, int* status,
                         double2* lu_out, int* indx_out, double* d_out, hipStream_t s,
                         const LuOpts& opts) {
  if (n < 1 || n > 64) return hipErrorInvalidValue;
  const long blocks = (long)ncfg * (minors ? n : 1);
  const size_t shm = (size_t)n * n * sizeof(double2);
  hipLaunchKernelGGL(k_lu_det, dim3((unsigned)blocks), dim3(64), shm, s, M, n, ldm, ncfg,
                     minors, logabs, phase, argsum, status, lu_out, indx_out, d_out, opts);
  return hipGetLastError();
}

// ------------------------------------------------------------------ dense overlaps
// M[c][s][t] = sum_i conj(phi[s][i]) psi[c][t][i] (src/one_body/noarch/one_body_project.C:
// 14-27 for every (s, t), src/slater_det/slater_det.C:33-37).  One workgroup per (c, t)
// streams its orbital once and accumulates 16 sinks at a time in FP64; the sinks are shared
// by every workgroup and stay cache resident.
template <typename T>
__global__ __launch_bounds__(256) void k_overlap_dense(const C2<T>* __restrict__ phi,
                                                       const C2<T>* __restrict__ psi,
                                                       double2* M, int ns, int nt, long vol,
                                                       long cstride) {
  constexpr int CH = 16;
  __shared__ double red[CH][2][4];
  const long ct = blockIdx.x;  // c * nt + t
  const long c = ct / nt;
  const int t = (int)(ct - c * nt);
  const int s0 = blockIdx.y * CH;
  const C2<T>* p = psi + c * cstride + (long)t * vol;
  double ar[CH], ai[CH];
#pragma unroll
  for (int q = 0; q < CH; ++q) { ar[q] = 0.0; ai[q] = 0.0; }
  for (long i = threadIdx.x; i < vol; i += 256) {
    const C2<T> b = p[i];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      if (s0 + q < ns) {
        const C2<T> f = phi[(long)(s0 + q) * vol + i];
        ar[q] += (double)f.x * b.x + (double)f.y * b.y;
        ai[q] += (double)f.x * b.y - (double)f.y * b.x;
      }
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < CH; ++q) {
    double x = ar[q], y = ai[q];
    for (int off = 32; off >= 1; off >>= 1) {
      x += __shfl_xor(x, off);
      y += __shfl_xor(y, off);
    }
    if (lane == 0) { red[q][0][w] = x; red[q][1][w] = y; }
  }
  __syncthreads();
  if (threadIdx.x < CH && s0 + (int)threadIdx.x < ns) {
    const int q = threadIdx.x;
    const double x = red[q][0][0] + red[q][0][1] + red[q][0][2] + red[q][0][3];
    const double y = red[q][1][0] + red[q][1][1] + red[q][1][2] + red[q][1][3];
    M[(c * ns + s0 + q) * nt + t] = make_double2(x, y);
  }
}

// The pair-sink matrix of TwoBody::Run (src/two_body/noarch/two_body_run.C:16-43, for every
// (s, t) as src/slater_det/slater_det3.C:115-122 fills phi[t + N s]):
//   M[c][s][t] = sum_i w[i] psi1[c][s][i] psi2[c][t][opp[i]]   (no conjugation)
// FP64 products and sums; one workgroup per (c, t) and 16 rows: column t's weighted mirror
// image w[i] psi2_t[opp[i]] is formed once per element and meets the 16 rows.
template <typename T>
__global__ __launch_bounds__(256) void k_pair_dense(const C2<T>* __restrict__ psi1,
                                                    const C2<T>* __restrict__ psi2,
                                                    const double* __restrict__ w,
                                                    const int* __restrict__ opp, double2* M,
                                                    int n1, int n2, long vol) {
  constexpr int CH = 16;
  __shared__ double red[CH][2][4];
  const long ct = blockIdx.x;  // c * n2 + t
  const long c = ct / n2;
  const int t = (int)(ct - c * n2);
  const int s0 = blockIdx.y * CH;
  const C2<T>* b2 = psi2 + ct * vol;
  const C2<T>* a1 = psi1 + c * n1 * vol;
  double ar[CH], ai[CH];
#pragma unroll
  for (int q = 0; q < CH; ++q) { ar[q] = 0.0; ai[q] = 0.0; }
  for (long i = threadIdx.x; i < vol; i += 256) {
    C2<T> u;
    if (opp) {
      u = b2[opp[i]];
    } else {  // the conjugate partner's image: conj psi2_t at the same site
      u = b2[i];
      u.y = -u.y;
    }
    const double br = w[i] * (double)u.x, bi = w[i] * (double)u.y;
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      if (s0 + q < n1) {
        const C2<T> a = a1[(long)(s0 + q) * vol + i];
        ar[q] += (double)a.x * br - (double)a.y * bi;
        ai[q] += (double)a.x * bi + (double)a.y * br;
      }
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < CH; ++q) {
    double x = ar[q], y = ai[q];
    for (int off = 32; off >= 1; off >>= 1) {
      x += __shfl_xor(x, off);
      y += __shfl_xor(y, off);
    }
    if (lane == 0) { red[q][0][wv] = x; red[q][1][wv] = y; }
  }
  __syncthreads();
  if (threadIdx.x < CH && s0 + (int)threadIdx.x < n1) {
    const int q = threadIdx.x;
    const double x = red[q][0][0] + red[q][0][1] + red[q][0][2] + red[q][0][3];
    const double y = red[q][1][0] + red[q][1][1] + red[q][1][2] + red[q][1][3];
    M[(c * n1 + s0 + q) * n2 + t] = make_double2(x, y);
  }
}

hipError_t launch_pair_dense(int prec, const void* psi1, const void* psi2, const double* w,
                             const int* opp, double2* M, int n1, int n2, long vol, int ncfg,
                             hipStream_t s) {
  dim3 grid((unsigned)((long)ncfg * n2), (unsigned)((n1 + 15) / 16));
  if (prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_pair_dense<double>, grid, dim3(256), 0, s, (const C2<double>*)psi1,
                       (const C2<double>*)psi2, w, opp, M, n1, n2, vol);
  else
    hipLaunchKernelGGL(k_pair_dense<float>, grid, dim3(256), 0, s, (const C2<float>*)psi1,
                       (const C2<float>*)psi2, w, opp, M, n1, n2, vol);
  return hipGetLastError();
}

__global__ void k_pair_expand(const double2* G, const double2* H, const int* col_vol,
                              const int* col_conj, double2* M, int N, int nvo, long ncfg) {
  const long n = ncfg * N * N;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long c = i / ((long)N * N);
    const int r = (int)(i - c * N * N), s = r / N, t = r - (r / N) * N;
    const int js = col_conj[s], jt = col_conj[t];
    const double2 g = (js == jt ? G : H)[(c * nvo + col_vol[s]) * nvo + col_vol[t]];
    M[i] = js ? make_double2(g.x, -g.y) : g;
  }
}

hipError_t launch_pair_expand(const double2* G, const double2* H, const int* col_vol,
                              const int* col_conj, double2* M, int N, int nvo, long ncfg,
                              hipStream_t s) {
  const long n = ncfg * N * N;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_pair_expand, dim3((unsigned)blocks), dim3(256), 0, s, G, H, col_vol,
                     col_conj, M, N, nvo, ncfg);
  return hipGetLastError();
}

// M[c][s][t] of orbitals stored as nvo volumes with conjugate partners, from the overlaps of
// the sinks with the stored volumes (Mp, [c][s][nvo]) and of their mirrors J phi_s with the
// volumes v0 .. v0 + nm - 1 (Mm, [c][s][nm]): column t is volume col_vol[t], or the conjugate
// of its mirrored-sink overlap when col_conj[t]
__global__ void k_sink_expand(const double2* Mp, const double2* Mm, const int* col_vol,
                              const int* col_conj, double2* M, int N, int nvo, int nm, int v0,
                              long ncfg) {
  const long n = ncfg * N * N;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long c = i / ((long)N * N);
    const int r = (int)(i - c * N * N), s = r / N, t = r - (r / N) * N;
    if (col_conj[t]) {
      const double2 g = Mm[(c * N + s) * nm + col_vol[t] - v0];
      M[i] = make_double2(g.x, -g.y);
    } else {
      M[i] = Mp[(c * N + s) * nvo + col_vol[t]];
    }
  }
}

hipError_t launch_sink_expand(const double2* Mp, const double2* Mm, const int* col_vol,
                              const int* col_conj, double2* M, int N, int nvo, int nm, int v0,
                              long ncfg, hipStream_t s) {
  const long n = ncfg * N * N;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_sink_expand, dim3((unsigned)blocks), dim3(256), 0, s, Mp, Mm, col_vol,
                     col_conj, M, N, nvo, nm, v0, ncfg);
  return hipGetLastError();
}

hipError_t launch_overlap_dense(int prec, const void* phi, const void* psi, double2* M,
                                int ns, int nt, long vol, int ncfg, hipStream_t s,
                                long cstride) {
  if (cstride == 0) cstride = nt * vol;
  dim3 grid((unsigned)((long)ncfg * nt), (unsigned)((ns + 15) / 16));
  if (prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_overlap_dense<double>, grid, dim3(256), 0, s,
                       (const C2<double>*)phi, (const C2<double>*)psi, M, ns, nt, vol, cstride);
  else
    hipLaunchKernelGGL(k_overlap_dense<float>, grid, dim3(256), 0, s, (const C2<float>*)phi,
                       (const C2<float>*)psi, M, ns, nt, vol, cstride);
  return hipGetLastError();
}

}  // namespace nrc
