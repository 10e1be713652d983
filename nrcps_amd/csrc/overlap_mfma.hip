// overlap_mfma.hip -- dense overlaps M[c][s][t] = sum_i conj(phi_s(i)) psi_{c,t}(i) on the
// matrix cores, as split-BF16x3 products with FP32 accumulation (general sinks: the
// reference's OneBody::Project / Cuda::InnerProduct, one_body_project.cu:17-34,
// cuda_utils.cu:88-127, for any sink vectors, not only plane waves).
//
// Complex products as one real GEMM over K = 2V interleaved (re, im) values:
//   Re M[s][t] = phi'_s . psi'_t,  Im M[s][t] = phi'_s . psi''_t,
// with phi' = (re, im) interleaved, psi' = (re, im) and psi'' = (im, -re).  The B matrix has
// two columns per orbital (2t: psi', 2t+1: psi''), built in registers from one load.
//
// Split-BF16x3: a = a_hi + a_lo with a_hi = bf16(a), a_lo = bf16(a - a_hi); a.b is taken as
// a_hi.b_hi + a_hi.b_lo + a_lo.b_hi (the dropped a_lo.b_lo is ~2^-18 relative).  Sinks are
// split once (they are shared by every configuration); orbitals are split as they are read.
//
// v_mfma_f32_16x16x32_bf16: lane l supplies A row (l % 16) and B column (l % 16), eight K
// values each for the lane group l / 16; D row (l / 16) * 4 + r, column l % 16.  A and B use
// the same lane/slot -> k assignment, so the sum over k is the full dot product.
//
// Each workgroup (4 waves) takes one configuration and one chunk of K; every wave
// accumulates all (sink tile, column tile) pairs for its K-steps, the waves are summed in
// LDS and the chunk's partial tile written out; a second kernel sums the chunks in FP64.
#include <algorithm>

#include "nrc_ops.h"

namespace nrc {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int OV_MAXT = 4;    // sink tiles of 16 (ns <= 64)
constexpr int OV_KSTEP = 32;  // reals per MFMA
constexpr int OV_WAVES = 4;

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// split 8 floats into bf16 hi and lo parts
__device__ __forceinline__ void split8(const float* v, bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const __bf16 h = to_bf16(v[q]);
    hi[q] = h;
    lo[q] = to_bf16(v[q] - (float)h);
  }
}

// split 8 floats into three bf16 parts, v = p0 + p1 + p2 to the last bit of FP32 (each part
// holds the next 8 significant bits of the remainder; both subtractions are exact)
__device__ __forceinline__ void split8_3(const float* v, bf16x8& p0, bf16x8& p1, bf16x8& p2) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const __bf16 h = to_bf16(v[q]);
    const float r = v[q] - (float)h;
    const __bf16 m = to_bf16(r);
    p0[q] = h;
    p1[q] = m;
    p2[q] = to_bf16(r - (float)m);
  }
}

// sinks phi[s][V] complex FP32 -> `parts` bf16 planes [part][s][2V] (interleaved re, im):
// parts = 2: hi, lo; parts = 3: the three-way split of split8_3
__global__ void k_split_bf16(const float2* phi, __bf16* out, long n, int parts) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const float2 v = phi[i];
    const float x[2] = {v.x, v.y};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const __bf16 h = to_bf16(x[j]);
      const float r = x[j] - (float)h;
      const __bf16 m = to_bf16(r);
      out[2 * i + j] = h;
      out[2 * n + 2 * i + j] = m;
      if (parts == 3) out[4 * n + 2 * i + j] = to_bf16(r - (float)m);
    }
  }
}

hipError_t launch_split_bf16(const float2* phi, void* out, int ns, long vol, hipStream_t s,
                             int parts) {
  if (parts != 2 && parts != 3) return hipErrorInvalidValue;
  const long n = (long)ns * vol;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_split_bf16, dim3((unsigned)blocks), dim3(256), 0, s, phi, (__bf16*)out,
                     n, parts);
  return hipGetLastError();
}

// grid (kchunks, ncfg); partial[c][chunk][MT 16][NT 32] FP32, re in column 2t, im in 2t + 1.
//
// A lane of column tile b holds orbital t = 16 b + l % 16 and reads its four complex values
// once; B's real-part column is (psi_re, psi_im) and its imaginary-part column
// (psi_im, -psi_re), which in bf16 is the real-part column with the halves of every dword
// swapped and the upper sign flipped (the split of -x is minus the split of x), so one split
// serves both accumulators.  Sinks and orbitals go through buffer resources (uniform bases,
// one lane offset per tile, the K position as the scalar offset; rows past ns / nt are past
// the resource's end and read zero), and the next step's values are loaded before the
// current step's products.
//
// TERMS = 6 (split-BF16x6): sinks and orbital values are split three ways, p = p0 + p1 + p2,
// exact to the last bit of FP32, and a.b is taken as the six products of total order <= 2
// (a0 b0, a0 b1, a1 b0, a0 b2, a1 b1, a2 b0; the dropped ones are below 2^-24 relative), so a
// plane-wave sink is represented exactly and the overlap error is FP32 accumulation alone.
template <int MT, int NT, int TERMS>
__global__ __launch_bounds__(256, (MT * NT <= 9 ? 2 : 1)) void k_overlap_bf16x3(
    const __bf16* __restrict__ phi_hi, const __bf16* __restrict__ phi_lo,
    const __bf16* __restrict__ phi_lo2,
    const float2* __restrict__ psi, float* __restrict__ partial, int ns, int nt, long vol,
    long kchunk, long cstride) {
  static_assert(TERMS == 3 || TERMS == 6, "terms");
  __shared__ float red[OV_WAVES - 1][4][64];
  const int chunk = blockIdx.x, nchunks = gridDim.x;
  const long c = blockIdx.y;
  // the wave index is uniform: in an SGPR, so the K position can be the scalar offset
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int row = lane & 15, grp = lane >> 4;
  const long K = 2 * vol;
  const long k0 = (long)chunk * kchunk;
  const long k1 = k0 + kchunk < K ? k0 + kchunk : K;
  const rsrc_t rh = make_rsrc(phi_hi, (uint32_t)(ns * K * 2));
  const rsrc_t rl = make_rsrc(phi_lo, (uint32_t)(ns * K * 2));
  const rsrc_t rl2 = make_rsrc(phi_lo2, TERMS == 6 ? (uint32_t)(ns * K * 2) : 0u);
  const rsrc_t rb = make_rsrc(psi + c * cstride, (uint32_t)(nt * vol * 8));
  constexpr uint32_t kFar = 0x7FFFFFF0u;  // past any resource: loads return zero
  uint32_t oa[MT], ob[NT];
#pragma unroll
  for (int a = 0; a < MT; ++a) {
    const int s = a * 16 + row;
    oa[a] = s < ns ? (uint32_t)(s * K + 8 * grp) * 2u : kFar;
  }
#pragma unroll
  for (int b = 0; b < NT; ++b) {
    const int t = b * 16 + row;
    ob[b] = t < nt ? (uint32_t)(t * vol + 4 * grp) * 8u : kFar;
  }
  f32x4 accr[MT][NT], acci[MT][NT];
#pragma unroll
  for (int a = 0; a < MT; ++a)
#pragma unroll
    for (int b = 0; b < NT; ++b) {
      accr[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
      acci[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
    }
  using u32x4 = __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t;
  // the sink (hi, lo) and orbital values of one step; lanes past the chunk end read zero (the
  // rows are contiguous, so a read past K would be the next row's)
  struct Step {
    u32x4 ah[MT], al[MT], a2[TERMS == 6 ? MT : 1];
    float4 b0[NT], b1[NT];
  };
  auto load = [&](Step& st, long kb) {
    const uint32_t sa = (uint32_t)kb * 2u, sb = (uint32_t)kb * 4u;
    const bool in = kb + 8 * grp < k1;
#pragma unroll
    for (int a = 0; a < MT; ++a) {
      const uint32_t o = in ? oa[a] : kFar;
      st.ah[a] = bload<u32x4>(rh, o, sa);
      st.al[a] = bload<u32x4>(rl, o, sa);
      if constexpr (TERMS == 6) st.a2[a] = bload<u32x4>(rl2, o, sa);
    }
#pragma unroll
    for (int b = 0; b < NT; ++b) {
      const uint32_t o = in ? ob[b] : kFar;
      st.b0[b] = bload<float4>(rb, o, sb);
      st.b1[b] = bload<float4>(rb, o + 16u, sb);
    }
  };
  // tile by tile: the orbital values are split, then the products run term by term over the
  // sink tiles, so consecutive MFMAs accumulate into different tiles
  auto compute = [&](const Step& st) {
#pragma unroll
    for (int b = 0; b < NT; ++b) {
      const float v[8] = {st.b0[b].x, st.b0[b].y, st.b0[b].z, st.b0[b].w,
                          st.b1[b].x, st.b1[b].y, st.b1[b].z, st.b1[b].w};
      bf16x8 pr[3];  // orbital parts (the third only with TERMS = 6)
      if constexpr (TERMS == 6) split8_3(v, pr[0], pr[1], pr[2]);
      else split8(v, pr[0], pr[1]);
      // imaginary-part column: (im, -re) per element = dword halves swapped, upper sign flipped
      bf16x8 pi[3];
#pragma unroll
      for (int j = 0; j < (TERMS == 6 ? 3 : 2); ++j) {
        const u32x4 u = __builtin_bit_cast(u32x4, pr[j]);
        u32x4 w;
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = ((u[q] >> 16) | (u[q] << 16)) ^ 0x80000000u;
        pi[j] = __builtin_bit_cast(bf16x8, w);
      }
      // (sink part, orbital part) of each term: TERMS 3: (0,0) (0,1) (1,0); 6: also (0,2)
      // (1,1) (2,0)
      constexpr int TA[6] = {0, 0, 1, 0, 1, 2}, TB[6] = {0, 1, 0, 2, 1, 0};
#pragma unroll
      for (int term = 0; term < TERMS; ++term)
#pragma unroll
        for (int a = 0; a < MT; ++a) {
          const u32x4 araw = TA[term] == 0 ? st.ah[a] : TA[term] == 1 ? st.al[a] : st.a2[a % (TERMS == 6 ? MT : 1)];
          const bf16x8 av = __builtin_bit_cast(bf16x8, araw);
          const bf16x8 br = pr[TB[term]], bi = pi[TB[term]];
          accr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, br, accr[a][b], 0, 0, 0);
          acci[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bi, acci[a][b], 0, 0, 0);
        }
    }
  };
  // two buffers: the next step's loads go out before the current step's products
  constexpr long S = OV_WAVES * OV_KSTEP;
  long kb = k0 + (long)wave * OV_KSTEP;
  Step r0, r1;
  load(r0, kb);
  while (kb < k1) {
    load(r1, kb + S);
    compute(r0);
    kb += S;
    if (kb >= k1) break;
    load(r0, kb + S);
    compute(r1);
    kb += S;
  }
  // sum the waves tile by tile: waves 1..3 park a tile in LDS, wave 0 adds and stores it
  constexpr int ldp = NT * 32;
  float* out = partial + (c * nchunks + chunk) * (long)(MT * 16) * ldp;
#pragma unroll
  for (int a = 0; a < MT; ++a)
#pragma unroll
    for (int b = 0; b < NT; ++b)
#pragma unroll
      for (int part = 0; part < 2; ++part) {
        const f32x4 v = part ? acci[a][b] : accr[a][b];
        if (wave > 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r) red[wave - 1][r][lane] = v[r];
        }
        __syncthreads();
        if (wave == 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            float x = v[r];
#pragma unroll
            for (int w = 0; w < OV_WAVES - 1; ++w) x += red[w][r][lane];
            out[(long)(a * 16 + grp * 4 + r) * ldp + 2 * (b * 16 + row) + part] = x;
          }
        }
        __syncthreads();
      }
}

// The pair-sink matrix M[c][s][t] = sum_i w[i] psi1[c][s][i] psi2[c][t][opp[i]] (TwoBody::Run,
// src/two_body/noarch/two_body_run.C:16-43) as the same split-BF16x3 GEMM: A rows are the
// configuration's own psi1 orbitals, split as they are read; a lane of column tile b holds
// orbital t = 16 b + l % 16 and builds both B columns from one load of b = w psi2_t(opp):
// (b_re, -b_im) for the real part and (b_im, b_re) for the imaginary part, accumulated
// separately.  The four mirrored sites of a lane's elements come as one int4; when they are
// consecutive (opp reverses each axis, so every group but the one holding z = 0 under
// periodic boundaries) the four values are two 16-byte loads (or 8 + 16 + 8 bytes from an
// odd site), reversed in registers; opp and wf must be 16-byte aligned.  The
// weights are rounded to FP32 as they are read.  Partial tiles keep re in column 2t and im
// in 2t + 1, as the overlap's.  CJ: B is w conj(psi2_t) at the same site (opp unused), the
// conjugate partner's image psi_{-t}(-p) = conj psi_t(p).
template <int MT, int NT, bool CJ>
__global__ __launch_bounds__(256) void k_pair_bf16x3(const float2* __restrict__ psi1,
                                                     const float2* __restrict__ psi2,
                                                     const double* __restrict__ wf,
                                                     const int* __restrict__ opp,
                                                     float* __restrict__ partial, int n1,
                                                     int n2, long vol, int mt, int nt,
                                                     long kchunk) {
  __shared__ float red[OV_WAVES - 1][4][64];
  const int chunk = blockIdx.x, nchunks = gridDim.x;
  const long c = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, grp = lane >> 4;
  const long K = 2 * vol;
  const long k0 = (long)chunk * kchunk;
  const long k1 = k0 + kchunk < K ? k0 + kchunk : K;
  f32x4 accr[MT][NT], acci[MT][NT];
#pragma unroll
  for (int a = 0; a < MT; ++a)
#pragma unroll
    for (int b = 0; b < NT; ++b) {
      accr[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
      acci[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
    }
  const float2* a1 = psi1 + c * n1 * vol;
  const float2* b2 = psi2 + c * n2 * vol;
  for (long kb = k0 + (long)wave * OV_KSTEP; kb < k1; kb += OV_WAVES * OV_KSTEP) {
    const long ka = kb + 8 * grp;  // this lane's 8 reals: complex elements j0 .. j0 + 3
    const long j0 = ka / 2;
    const bool in = ka < K;
    bf16x8 ah[MT], al[MT];
#pragma unroll
    for (int a = 0; a < MT; ++a) {
      if (a < mt) {
        const int s = a * 16 + row;
        float v[8];
        if (s < n1 && in) {
          const float4* p = reinterpret_cast<const float4*>(a1 + (long)s * vol + j0);
          const float4 u0 = p[0], u1 = p[1];
          v[0] = u0.x; v[1] = u0.y; v[2] = u0.z; v[3] = u0.w;
          v[4] = u1.x; v[5] = u1.y; v[6] = u1.z; v[7] = u1.w;
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q) v[q] = 0.f;
        }
        split8(v, ah[a], al[a]);
      }
    }
    int4 st = make_int4(0, 0, 0, 0);
    float wt[4] = {0.f, 0.f, 0.f, 0.f};
    if (in) {
      if constexpr (!CJ) st = *reinterpret_cast<const int4*>(opp + j0);
      const double2* wp = reinterpret_cast<const double2*>(wf + j0);
      const double2 w01 = wp[0], w23 = wp[1];
      wt[0] = (float)w01.x; wt[1] = (float)w01.y; wt[2] = (float)w23.x; wt[3] = (float)w23.y;
    }
    // consecutive: the mirrored sites are st.w .. st.w + 3 for elements 3, 2, 1, 0
    const bool run = st.x == st.w + 3 && st.y == st.w + 2 && st.z == st.w + 1;
    const bool even = (st.w & 1) == 0;
#pragma unroll
    for (int b = 0; b < NT; ++b) {
      const int t = b * 16 + row;
      float vr[8], vi[8];
      if (b < nt && t < n2 && in) {
        const float2* p = b2 + (long)t * vol;
        float2 u[4];
        if constexpr (CJ) {  // the same four sites, conjugated
          const float4* q = reinterpret_cast<const float4*>(p + j0);
          const float4 u01 = q[0], u23 = q[1];
          u[0] = make_float2(u01.x, -u01.y);
          u[1] = make_float2(u01.z, -u01.w);
          u[2] = make_float2(u23.x, -u23.y);
          u[3] = make_float2(u23.z, -u23.w);
        } else if (run && even) {  // 16-byte aligned
          const float4* q = reinterpret_cast<const float4*>(p + st.w);
          const float4 u32 = q[0], u10 = q[1];
          u[3] = make_float2(u32.x, u32.y);
          u[2] = make_float2(u32.z, u32.w);
          u[1] = make_float2(u10.x, u10.y);
          u[0] = make_float2(u10.z, u10.w);
        } else if (run) {  // odd start: 8 + 16 (aligned) + 8 bytes
          const float4 u21 = *reinterpret_cast<const float4*>(p + st.w + 1);
          u[3] = p[st.w];
          u[2] = make_float2(u21.x, u21.y);
          u[1] = make_float2(u21.z, u21.w);
          u[0] = p[st.w + 3];
        } else {
          u[0] = p[st.x];
          u[1] = p[st.y];
          u[2] = p[st.z];
          u[3] = p[st.w];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float br = wt[q] * u[q].x, bi = wt[q] * u[q].y;
          vr[2 * q] = br;
          vr[2 * q + 1] = -bi;
          vi[2 * q] = bi;
          vi[2 * q + 1] = br;
        }
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) vr[q] = vi[q] = 0.f;
      }
      if (b < nt) {
        bf16x8 rh, rl, ih, il;
        split8(vr, rh, rl);
        split8(vi, ih, il);
#pragma unroll
        for (int a = 0; a < MT; ++a) {
          if (a < mt) {
            accr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], rh, accr[a][b], 0, 0, 0);
            accr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], rl, accr[a][b], 0, 0, 0);
            accr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], rh, accr[a][b], 0, 0, 0);
            acci[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], ih, acci[a][b], 0, 0, 0);
            acci[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], il, acci[a][b], 0, 0, 0);
            acci[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], ih, acci[a][b], 0, 0, 0);
          }
        }
      }
    }
  }
  // sum the waves tile by tile; column 2t holds re, 2t + 1 im
  const int ldp = nt * 32;
  float* out = partial + (c * nchunks + chunk) * (long)(mt * 16) * ldp;
#pragma unroll
  for (int a = 0; a < MT; ++a)
#pragma unroll
    for (int b = 0; b < NT; ++b)
#pragma unroll
      for (int part = 0; part < 2; ++part) {
        if (a < mt && b < nt) {  // uniform over the workgroup
          const f32x4 v = part ? acci[a][b] : accr[a][b];
          if (wave > 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) red[wave - 1][r][lane] = v[r];
          }
          __syncthreads();
          if (wave == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              float x = v[r];
#pragma unroll
              for (int w = 0; w < OV_WAVES - 1; ++w) x += red[w][r][lane];
              out[(long)(a * 16 + grp * 4 + r) * ldp + 2 * (b * 16 + row) + part] = x;
            }
          }
          __syncthreads();
        }
      }
}

// The two pair products of conjugate partners in one pass over the stored volumes psi[c][n]:
// G[c][s][t] = sum_i w[i] psi_s(i) psi_t(opp i) and H[c][s][t] = sum_i w[i] psi_s(i)
// conj psi_t(i).  A lane of A tile a holds orbital s = 16 a + l % 16 at the step's four
// sites, which are also H's B values for the orbital t = s of tile b = a, so every step reads
// the sites once and their mirrors once.  The imaginary-part B column (bi, br) is the
// real-part one (br, -bi) with the halves of every dword swapped and the lower sign flipped,
// so one split serves both; the products run term by term over the sink tiles.  partial
// holds G's chunk tiles, then H's, in k_pair_bf16x3's layout.
template <int MT>
__global__ __launch_bounds__(256) void k_pair_gh_bf16x3(const float2* __restrict__ psi,
                                                        const double* __restrict__ wf,
                                                        const int* __restrict__ opp,
                                                        float* __restrict__ partial, int n,
                                                        long vol, long kchunk) {
  using u32x4 = __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t;
  __shared__ float red[OV_WAVES - 1][4][64];
  const int chunk = blockIdx.x, nchunks = gridDim.x;
  const long c = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, grp = lane >> 4;
  const long K = 2 * vol;
  const long k0 = (long)chunk * kchunk;
  const long k1 = k0 + kchunk < K ? k0 + kchunk : K;
  f32x4 acc[4][MT][MT];  // G re, G im, H re, H im
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
      for (int b = 0; b < MT; ++b) acc[m][a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const float2* base = psi + c * n * vol;
  for (long kb = k0 + (long)wave * OV_KSTEP; kb < k1; kb += OV_WAVES * OV_KSTEP) {
    const long ka = kb + 8 * grp;  // this lane's 8 reals: complex elements j0 .. j0 + 3
    const long j0 = ka / 2;
    const bool in = ka < K;
    float va[MT][8];
    bf16x8 ah[MT], al[MT];
#pragma unroll
    for (int a = 0; a < MT; ++a) {
      const int s = a * 16 + row;
      if (s < n && in) {
        const float4* p = reinterpret_cast<const float4*>(base + (long)s * vol + j0);
        const float4 u0 = p[0], u1 = p[1];
        va[a][0] = u0.x; va[a][1] = u0.y; va[a][2] = u0.z; va[a][3] = u0.w;
        va[a][4] = u1.x; va[a][5] = u1.y; va[a][6] = u1.z; va[a][7] = u1.w;
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) va[a][q] = 0.f;
      }
      split8(va[a], ah[a], al[a]);
    }
    int4 st = make_int4(0, 0, 0, 0);
    float wt[4] = {0.f, 0.f, 0.f, 0.f};
    if (in) {
      st = *reinterpret_cast<const int4*>(opp + j0);
      const double2* wp = reinterpret_cast<const double2*>(wf + j0);
      const double2 w01 = wp[0], w23 = wp[1];
      wt[0] = (float)w01.x; wt[1] = (float)w01.y; wt[2] = (float)w23.x; wt[3] = (float)w23.y;
    }
    // consecutive: the mirrored sites are st.w .. st.w + 3 for elements 3, 2, 1, 0
    const bool run = st.x == st.w + 3 && st.y == st.w + 2 && st.z == st.w + 1;
    const bool even = (st.w & 1) == 0;
#pragma unroll
    for (int b = 0; b < MT; ++b) {
      const int t = b * 16 + row;
#pragma unroll
      for (int m = 0; m < 2; ++m) {  // G: the mirrored sites; H: the same sites conjugated
        float2 u[4];
        if (m == 1) {
#pragma unroll
          for (int q = 0; q < 4; ++q) u[q] = make_float2(va[b][2 * q], -va[b][2 * q + 1]);
        } else if (t < n && in) {
          const float2* p = base + (long)t * vol;
          if (run && even) {  // 16-byte aligned
            const float4* q = reinterpret_cast<const float4*>(p + st.w);
            const float4 u32 = q[0], u10 = q[1];
            u[3] = make_float2(u32.x, u32.y);
            u[2] = make_float2(u32.z, u32.w);
            u[1] = make_float2(u10.x, u10.y);
            u[0] = make_float2(u10.z, u10.w);
          } else if (run) {  // odd start: 8 + 16 (aligned) + 8 bytes
            const float4 u21 = *reinterpret_cast<const float4*>(p + st.w + 1);
            u[3] = p[st.w];
            u[2] = make_float2(u21.x, u21.y);
            u[1] = make_float2(u21.z, u21.w);
            u[0] = p[st.w + 3];
          } else {
            u[0] = p[st.x];
            u[1] = p[st.y];
            u[2] = p[st.z];
            u[3] = p[st.w];
          }
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) u[q] = make_float2(0.f, 0.f);
        }
        float vr[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          vr[2 * q] = wt[q] * u[q].x;
          vr[2 * q + 1] = -(wt[q] * u[q].y);
        }
        bf16x8 rh, rl;
        split8(vr, rh, rl);
        const u32x4 uh = __builtin_bit_cast(u32x4, rh), ul = __builtin_bit_cast(u32x4, rl);
        u32x4 vh, vl;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          vh[q] = ((uh[q] >> 16) | (uh[q] << 16)) ^ 0x00008000u;
          vl[q] = ((ul[q] >> 16) | (ul[q] << 16)) ^ 0x00008000u;
        }
        const bf16x8 ih = __builtin_bit_cast(bf16x8, vh), il = __builtin_bit_cast(bf16x8, vl);
#pragma unroll
        for (int term = 0; term < 3; ++term)
#pragma unroll
          for (int a = 0; a < MT; ++a) {
            const bf16x8 av = term == 2 ? al[a] : ah[a];
            const bf16x8 br = term == 1 ? rl : rh, bi = term == 1 ? il : ih;
            acc[2 * m][a][b] =
                __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, br, acc[2 * m][a][b], 0, 0, 0);
            acc[2 * m + 1][a][b] =
                __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bi, acc[2 * m + 1][a][b], 0, 0, 0);
          }
      }
    }
  }
  // sum the waves tile by tile; column 2t holds re, 2t + 1 im
  constexpr int ldp = MT * 32;
  const long half = (long)gridDim.y * nchunks * (MT * 16) * ldp;
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
      for (int b = 0; b < MT; ++b)
#pragma unroll
        for (int part = 0; part < 2; ++part) {
          const f32x4 v = acc[2 * m + part][a][b];
          if (wave > 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) red[wave - 1][r][lane] = v[r];
          }
          __syncthreads();
          if (wave == 0) {
            float* out = partial + m * half + (c * nchunks + chunk) * (long)(MT * 16) * ldp;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              float x = v[r];
#pragma unroll
              for (int w = 0; w < OV_WAVES - 1; ++w) x += red[w][r][lane];
              out[(long)(a * 16 + grp * 4 + r) * ldp + 2 * (b * 16 + row) + part] = x;
            }
          }
          __syncthreads();
        }
}

// k_pair_gh_bf16x3 for opp(p) = -p on a cubic lattice (the conjugate partners' case), with
// every orbital value read from HBM once and split once.  The workgroups take units of the
// lattice: the plane pair (x, L - x) for x = 1 .. L/2 - 1 -- the sites p of plane x and their
// mirrors in plane L - x, each loaded pair giving the terms of p and of -p -- and the planes 0
// and L/2, mirrored onto themselves, whose sites each give their own terms (the mirror read
// hits the cache: the same workgroup reads the plane).  A lane's four consecutive sites (y,
// z .. z + 3), z a multiple of 4, have the mirrors (L - y, L - z .. L - z - 3): one reversed
// run, or for z = 0 the site (L - y, 0) and a run of three.
//
// w is even (w(-p) = w(p)), so with u(p) = sqrt|w(p)| psi(p) and sg(p) the sign of w(p)
//   G[s][t] = sum_p sg(p) u_s(p) u_t(-p),   H[s][t] = sum_p sg(p) u_s(p) conj u_t(p):
// the scaled values of every stored volume, split once into bf16 (hi, lo), are both the A and
// the B operands.  In the interleaved (re, im) layout, with A = (a_r, a_i) and B = (b_r, b_i):
//   Re a b      = A . (b_r, -b_i)        Im a b      = swap(A) . B
//   Re a conj b = A . B                  Im a conj b = swap(A) . (b_r, -b_i)
// (swap exchanges the halves of every dword; the split of -x is minus the split of x), so a
// step's operands take one dword rotate per A dword and one sign xor per B dword, the sign of
// w riding on the B side (ws holds sg(p) sqrt|w(p)| in FP32).
//
// G is symmetric and H Hermitian.  So over a plane pair the terms of the mirrored sites add
// the transpose of the sites' own G terms, and only H's tiles on and above the diagonal are
// formed: the sites of plane x give G (every tile) and H (upper tiles), their mirrors H (upper
// tiles) -- 10 tile products a step at MT = 2 instead of 16.  k_pair_gh_reduce assembles G
// (adding the transpose for plane pairs) and H (conjugating the upper tiles into the lower).
// Partial tiles per unit in k_pair_gh_bf16x3's layout; H's tiles below the diagonal are not
// written.
template <int MT, bool NEG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MT == 2 ? 3 : 1)))
void k_pair_gh_mirror(const float2* __restrict__ psi,
                                                        const float* __restrict__ ws,
                                                        float* __restrict__ partial, int n,
                                                        int L) {
  using u32x4 = __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t;
  __shared__ float red[OV_WAVES - 1][4][64];
  const int unit = blockIdx.x, nunits = gridDim.x;
  const long c = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, grp = lane >> 4;
  const int P = L * L;
  const long V = (long)L * P;
  const int x = unit < L / 2 - 1 ? unit + 1 : (unit == L / 2 - 1 ? 0 : L / 2);
  const int xm = (L - x) % L;
  const bool pair = x != xm;
  f32x4 acc[4][MT][MT];  // G re, G im, H re, H im
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
      for (int b = 0; b < MT; ++b) acc[m][a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const float2* base = psi + c * n * V;
  // (loading the next step's values ahead of the products needs registers that cost the
  // kernel its third wave per SIMD, and was slower)
  for (int i0 = wave * 16; i0 < P; i0 += OV_WAVES * 16) {
    const int i = i0 + 4 * grp;  // this lane's four sites of the plane
    const int y = i / L, z = i - y * L;
    const long p0 = (long)x * P + i;
    const long mrow = (long)xm * P + (long)((L - y) % L) * L;
    float va[MT][8], vm[MT][8];
#pragma unroll
    for (int a = 0; a < MT; ++a) {
      const int s = a * 16 + row;
      if (s < n) {
        const float2* o = base + (long)s * V;
        const float4* q = reinterpret_cast<const float4*>(o + p0);
        const float4 u0 = q[0], u1 = q[1];
        va[a][0] = u0.x; va[a][1] = u0.y; va[a][2] = u0.z; va[a][3] = u0.w;
        va[a][4] = u1.x; va[a][5] = u1.y; va[a][6] = u1.z; va[a][7] = u1.w;
        float2 m[4];
        if (z > 0) {  // mirrors of z .. z + 3: m0 + 3 .. m0, m0 = L - z - 3 (odd)
          const long m0 = mrow + (L - z - 3);
          const float4 u21 = *reinterpret_cast<const float4*>(o + m0 + 1);
          m[3] = o[m0];
          m[2] = make_float2(u21.x, u21.y);
          m[1] = make_float2(u21.z, u21.w);
          m[0] = o[m0 + 3];
        } else {
          m[0] = o[mrow];
          m[1] = o[mrow + L - 1];
          m[2] = o[mrow + L - 2];
          m[3] = o[mrow + L - 3];
        }
#pragma unroll
        for (int q2 = 0; q2 < 4; ++q2) {
          vm[a][2 * q2] = m[q2].x;
          vm[a][2 * q2 + 1] = m[q2].y;
        }
      } else {
#pragma unroll
        for (int q2 = 0; q2 < 8; ++q2) va[a][q2] = vm[a][q2] = 0.f;
      }
    }
    const float4 wv = *reinterpret_cast<const float4*>(ws + p0);
    const float wa[4] = {fabsf(wv.x), fabsf(wv.y), fabsf(wv.z), fabsf(wv.w)};
    // per site dword: m1 = the sign of w on both halves, m2 = m1 with the imaginary sign flipped
    u32x4 m1 = {0u, 0u, 0u, 0u}, m2 = {0x80000000u, 0x80000000u, 0x80000000u, 0x80000000u};
    if constexpr (NEG) {
      const float wq[4] = {wv.x, wv.y, wv.z, wv.w};
#pragma unroll
      for (int q2 = 0; q2 < 4; ++q2) {
        const uint32_t sb = __float_as_uint(wq[q2]) & 0x80000000u;
        m1[q2] = sb | (sb >> 16);
        m2[q2] = m1[q2] ^ 0x80000000u;
      }
    }
    // scaled values of the sites (p) and of their mirrors (m), split into hi and lo
    u32x4 ph[MT], pl[MT], mh[MT], ml[MT];
#pragma unroll
    for (int a = 0; a < MT; ++a) {
      float t[8];
      bf16x8 h, l;
#pragma unroll
      for (int q2 = 0; q2 < 8; ++q2) t[q2] = va[a][q2] * wa[q2 >> 1];
      split8(t, h, l);
      ph[a] = __builtin_bit_cast(u32x4, h);
      pl[a] = __builtin_bit_cast(u32x4, l);
#pragma unroll
      for (int q2 = 0; q2 < 8; ++q2) t[q2] = vm[a][q2] * wa[q2 >> 1];
      split8(t, h, l);
      mh[a] = __builtin_bit_cast(u32x4, h);
      ml[a] = __builtin_bit_cast(u32x4, l);
    }
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      if (part == 1 && !pair) break;
      // part 0, the terms of the sites p: A = u(p), G's B = u(-p), H's B = u(p);
      // part 1, the mirrored sites' H terms: A = B = u(-p) (their G terms are the transpose)
      const u32x4* Ah = part ? mh : ph;
      const u32x4* Al = part ? ml : pl;
      bf16x8 a_h[MT], a_l[MT], s_h[MT], s_l[MT];  // A and swap(A)
#pragma unroll
      for (int a = 0; a < MT; ++a) {
        u32x4 xh, xl;
#pragma unroll
        for (int q2 = 0; q2 < 4; ++q2) {
          xh[q2] = (Ah[a][q2] >> 16) | (Ah[a][q2] << 16);
          xl[q2] = (Al[a][q2] >> 16) | (Al[a][q2] << 16);
        }
        a_h[a] = __builtin_bit_cast(bf16x8, Ah[a]);
        a_l[a] = __builtin_bit_cast(bf16x8, Al[a]);
        s_h[a] = __builtin_bit_cast(bf16x8, xh);
        s_l[a] = __builtin_bit_cast(bf16x8, xl);
      }
#pragma unroll
      for (int b = 0; b < MT; ++b) {
        // G: Re = A . (B ^ m2), Im = swap(A) . (B ^ m1), B = u(-p);
        // H: Re = A . (B ^ m1), Im = swap(A) . (B ^ m2), B = A's values
        bf16x8 gr[2], gi[2], hr[2], hi[2];
#pragma unroll
        for (int lo = 0; lo < 2; ++lo) {
          const u32x4 gb = lo ? ml[b] : mh[b], hb = lo ? Al[b] : Ah[b];
          u32x4 t0, t1, t2, t3;
#pragma unroll
          for (int q2 = 0; q2 < 4; ++q2) {
            t0[q2] = gb[q2] ^ m2[q2];
            t1[q2] = NEG ? gb[q2] ^ m1[q2] : gb[q2];
            t2[q2] = NEG ? hb[q2] ^ m1[q2] : hb[q2];
            t3[q2] = hb[q2] ^ m2[q2];
          }
          gr[lo] = __builtin_bit_cast(bf16x8, t0);
          gi[lo] = __builtin_bit_cast(bf16x8, t1);
          hr[lo] = __builtin_bit_cast(bf16x8, t2);
          hi[lo] = __builtin_bit_cast(bf16x8, t3);
        }
#pragma unroll
        for (int term = 0; term < 3; ++term) {  // (hi, hi), (hi, lo), (lo, hi)
          const int tb = term == 1 ? 1 : 0;
#pragma unroll
          for (int a = 0; a < MT; ++a) {
            const bf16x8 av = term == 2 ? a_l[a] : a_h[a];
            const bf16x8 sv = term == 2 ? s_l[a] : s_h[a];
            if (part == 0) {
              acc[0][a][b] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, gr[tb], acc[0][a][b], 0, 0, 0);
              acc[1][a][b] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(sv, gi[tb], acc[1][a][b], 0, 0, 0);
            }
            if (a <= b) {
              acc[2][a][b] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, hr[tb], acc[2][a][b], 0, 0, 0);
              acc[3][a][b] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(sv, hi[tb], acc[3][a][b], 0, 0, 0);
            }
          }
        }
      }
    }
  }
  constexpr int ldp = MT * 32;
  const long half = (long)gridDim.y * nunits * (MT * 16) * ldp;
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
      for (int b = 0; b < MT; ++b)
#pragma unroll
        for (int part = 0; part < 2; ++part) {
          if (m == 1 && a > b) continue;  // H below the diagonal: the reduce conjugates
          const f32x4 v = acc[2 * m + part][a][b];
          if (wave > 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) red[wave - 1][r][lane] = v[r];
          }
          __syncthreads();
          if (wave == 0) {
            float* out = partial + m * half + (c * nunits + unit) * (long)(MT * 16) * ldp;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              float xs = v[r];
#pragma unroll
              for (int w = 0; w < OV_WAVES - 1; ++w) xs += red[w][r][lane];
              out[(long)(a * 16 + grp * 4 + r) * ldp + 2 * (b * 16 + row) + part] = xs;
            }
          }
          __syncthreads();
        }
}

// G and H of the mirror form from its unit partials (FP64 sums): G[s][t] over every unit, plus
// G[t][s] over the plane pairs (units below npair); H[s][t] from the tiles on and above the
// diagonal, conj H[t][s] below it
__global__ void k_pair_gh_reduce(const float* partial, double2* G, double2* H, int n, int nunits,
                                 int npair, int mt, long ncfg) {
  const long total = ncfg * n * n;
  const int ldp = mt * 32;
  const long tile = (long)(mt * 16) * ldp;
  const long half = ncfg * nunits * tile;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const long c = i / ((long)n * n);
    const int r = (int)(i - c * n * n), s = r / n, t = r - (r / n) * n;
    double gr = 0.0, gi = 0.0, hr = 0.0, hi = 0.0;
    const bool up = s / 16 <= t / 16;
    const int hs = up ? s : t, ht = up ? t : s;
    for (int u = 0; u < nunits; ++u) {
      const float* p = partial + (c * nunits + u) * tile;
      gr += (double)p[(long)s * ldp + 2 * t];
      gi += (double)p[(long)s * ldp + 2 * t + 1];
      if (u < npair) {
        gr += (double)p[(long)t * ldp + 2 * s];
        gi += (double)p[(long)t * ldp + 2 * s + 1];
      }
      hr += (double)p[half + (long)hs * ldp + 2 * ht];
      hi += (double)p[half + (long)hs * ldp + 2 * ht + 1];
    }
    G[i] = make_double2(gr, gi);
    H[i] = make_double2(hr, up ? hi : -hi);
  }
}

// M[c][s][t] = sum over chunks (FP64): re from column 2t, im from column 2t + 1
__global__ void k_overlap_reduce(const float* partial, double2* M, int ns, int nt, int nchunks,
                                 int mt, int nc_t, long ncfg) {
  const long n = ncfg * ns * nt;
  const int ldp = nc_t * 16;
  const long tile = (long)(mt * 16) * ldp;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long c = i / ((long)ns * nt);
    const int r = (int)(i - c * ns * nt), s = r / nt, t = r - (r / nt) * nt;
    double re = 0.0, im = 0.0;
    for (int k = 0; k < nchunks; ++k) {
      const float* p = partial + (c * nchunks + k) * tile + (long)s * ldp;
      re += (double)p[2 * t];
      im += (double)p[2 * t + 1];
    }
    M[i] = make_double2(re, im);
  }
}

// chunk of K per workgroup: a multiple of the 4-wave step, about 16 chunks per configuration
long overlap_bf16x3_kchunk(long vol) {
  const long K = 2 * vol, step = OV_WAVES * OV_KSTEP;
  long kc = (K / 16 + step - 1) / step * step;
  return kc < step ? step : kc;
}

size_t overlap_bf16x3_partial_floats(int ns, int nt, long vol, int ncfg) {
  const int mt = (ns + 15) / 16, ntt = (nt + 15) / 16;
  const long K = 2 * vol, kc = overlap_bf16x3_kchunk(vol);
  const long nchunks = (K + kc - 1) / kc;
  return (size_t)ncfg * nchunks * (mt * 16) * (ntt * 32);
}

hipError_t launch_overlap_bf16x3(const void* phi_split, const float2* psi, double2* M, int ns,
                                 int nt, long vol, int ncfg, float* partial, hipStream_t s,
                                 int terms, long cstride) {
  if (terms != 3 && terms != 6) return hipErrorInvalidValue;
  if (cstride == 0) cstride = nt * vol;
  if (ns < 1 || ns > 16 * OV_MAXT || nt < 1 || nt > 16 * OV_MAXT || (vol & 3))
    return hipErrorInvalidValue;
  const int mt = (ns + 15) / 16, ntt = (nt + 15) / 16;
  const long K = 2 * vol, kc = overlap_bf16x3_kchunk(vol);
  const int nchunks = (int)((K + kc - 1) / kc);
  const __bf16* hi = (const __bf16*)phi_split;
  const __bf16* lo = hi + 2 * (long)ns * vol;
  const __bf16* lo2 = lo + 2 * (long)ns * vol;
  const dim3 grid(nchunks, ncfg), block(64 * OV_WAVES);
  // one instantiation per (sink tiles, orbital tiles), so only the used accumulators exist
#define NRC_OV_CASE(A, B)                                                                    \
  if (mt == A && ntt == B && terms == 3) {                                                   \
    hipLaunchKernelGGL((k_overlap_bf16x3<A, B, 3>), grid, block, 0, s, hi, lo, lo, psi,      \
                       partial, ns, nt, vol, kc, cstride);                                   \
  } else if (mt == A && ntt == B) {                                                          \
    hipLaunchKernelGGL((k_overlap_bf16x3<A, B, 6>), grid, block, 0, s, hi, lo, lo2, psi,     \
                       partial, ns, nt, vol, kc, cstride);                                   \
  }
#define NRC_OV_ROW(A) NRC_OV_CASE(A, 1) NRC_OV_CASE(A, 2) NRC_OV_CASE(A, 3) NRC_OV_CASE(A, 4)
  NRC_OV_ROW(1) NRC_OV_ROW(2) NRC_OV_ROW(3) NRC_OV_ROW(4)
#undef NRC_OV_ROW
#undef NRC_OV_CASE
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long n = (long)ncfg * ns * nt;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  // column pairs (2t, 2t + 1) over 2 ntt tiles of 16
  hipLaunchKernelGGL(k_overlap_reduce, dim3((unsigned)blocks), dim3(256), 0, s, partial, M, ns,
                     nt, nchunks, mt, 2 * ntt, (long)ncfg);
  return hipGetLastError();
}

// ------------------------------------------------------------------ overlaps, sink-row form
// The same products with the real / imaginary split carried by the sinks instead of the
// orbitals: A has two rows per sink, row 2s = (phi_re, phi_im) and row 2s + 1 = (-phi_im,
// phi_re) interleaved along K, so that with B the plain orbital values (psi_re, psi_im),
// D[2s][t] = Re conj(phi_s) . psi_t and D[2s + 1][t] = Im.  The sinks are split once per call
// (they are shared by every configuration); the orbital side needs no per-element swap, and
// 2 ns rows pack into fewer tiles of 16 than ns rows with two column sets (33 sinks x 33
// orbitals: 5 x 3 tiles instead of 3 x 3 x 2).  Each thread splits its eight orbital values
// with packed conversions: h = bf16x2(x0, x1) once, the floats of h by shifts, the remainders
// in FP32, the next part packed again -- bit for bit the parts of split8 / split8_3.

typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));

// x0, x1 -> packed bf16 (RNE) and the FP32 remainders x - float(bf16(x))
__device__ __forceinline__ uint32_t split_pk(float x0, float x1, float& r0, float& r1) {
  const uint32_t h =
      __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2v){x0, x1}, bf16x2v));
  r0 = x0 - __builtin_bit_cast(float, h << 16);
  r1 = x1 - __builtin_bit_cast(float, h & 0xFFFF0000u);
  return h;
}

template <int NP>
__device__ __forceinline__ void split8_pk(const float* v, bf16x8* p) {
  using u32x4 = __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t;
  u32x4 w[NP];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    float r0 = v[2 * q], r1 = v[2 * q + 1];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const float a0 = r0, a1 = r1;
      w[j][q] = split_pk(a0, a1, r0, r1);
    }
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) p[j] = __builtin_bit_cast(bf16x8, w[j]);
}

// sinks phi[s][V] -> `parts` planes [part][2 ns][2V]: row 2s = (re, im), row 2s + 1 = (-im, re)
__global__ void k_split_rows(const float2* phi, __bf16* out, long n, long vol, int parts) {
  const long K = 2 * vol, plane = 2 * n * 2;  // n = ns V elements; a plane holds 2 ns rows of K
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long sidx = i / vol, site = i - sidx * vol;
    const float2 v = phi[i];
    const float x[4] = {v.x, v.y, -v.y, v.x};  // row 2s: (re, im); row 2s + 1: (-im, re)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long o = (2 * sidx + (j >> 1)) * K + 2 * site + (j & 1);
      const __bf16 h = to_bf16(x[j]);
      const float r = x[j] - (float)h;
      const __bf16 m = to_bf16(r);
      out[o] = h;
      out[plane + o] = m;
      if (parts == 3) out[2 * plane + o] = to_bf16(r - (float)m);
    }
  }
}

// M[c][s][t] = sum over chunks (FP64) of rows 2s (re) and 2s + 1 (im); column (c, t) is
// column c ntp + t of the tiles of ntt 16 columns
__global__ void k_overlap_rows_reduce(const float* partial, double2* M, int ns, int nt,
                                      long ntp, int nchunks, int mr, int ntt, long ncfg) {
  const long n = ncfg * ns * nt;
  const int ldp = ntt * 16;
  const long tile = (long)(mr * 16) * ldp;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long c = i / ((long)ns * nt);
    const int r = (int)(i - c * ns * nt), s = r / nt, t = r - (r / nt) * nt;
    const long col = c * ntp + t, tl = col / ldp;
    const int j = (int)(col - tl * ldp);
    double re = 0.0, im = 0.0;
    for (int k = 0; k < nchunks; ++k) {
      const float* p = partial + (tl * nchunks + k) * tile + (long)(2 * s) * ldp + j;
      re += (double)p[0];
      im += (double)p[ldp];
    }
    M[i] = make_double2(re, im);
  }
}

// Column tiling of the sink-row form.  The columns (c, t) of all configurations form one
// sequence, column c nt + t, cut into tiles of 16 NT columns, so that no tile carries padded
// columns when nt is not a multiple of 16 (17 stored volumes per configuration of the
// partnered 33-orbital benchmark would otherwise fill 17 of 32 columns).  A tile's orbitals
// are addressed from one buffer resource at its first configuration; where that range could
// pass 2^31 bytes (or for one configuration) every configuration starts a tile of its own
// (ntp = 16 NT columns per configuration).
struct OvlCols {
  int nt_tiles;  // NT: 16-column tiles per wave
  long ntp;      // column stride of a configuration
  long ntiles;   // tiles of 16 NT columns
};
// shared tiles of 48 columns (three 16-column tiles per wave) where the row tiles leave the
// registers for them (up to 7 row tiles, 112 sink rows), else 32; wider for nt > 48
static int ovl_flat_nt(int nt, int mr) {
  const int ntt = (nt + 15) / 16;
  return ntt >= 3 ? ntt : (mr <= 7 ? 3 : 2);
}
static OvlCols ovl_cols(int nt, int mr, long vol, long ncfg, long cstride) {
  const int ntt = (nt + 15) / 16, ntf = ovl_flat_nt(nt, mr);
  const long span = (16L * ntf - 1) / nt + 2;  // configurations a tile can reach
  if (ncfg > 1 && ((span - 1) * cstride + (long)nt * vol) * 8 <= 0x7FFF0000L) {
    const long ncol = ncfg * nt, w = 16L * ntf;
    return {ntf, (long)nt, (ncol + w - 1) / w};
  }
  return {ntt, 16L * ntt, ncfg};
}

// partial tiles of the sink-row form: [tile][chunk][2 ns rows padded to 16][16 NT], sized
// for either tiling (the shared tiles hold at most ncfg nt + 63 columns) and non-decreasing in
// nt, so a workspace sized for nt columns serves any smaller nt
size_t overlap_rows_partial_floats(int ns, int nt, long vol, int ncfg) {
  const int mr = (2 * ns + 15) / 16, ntt = (nt + 15) / 16;
  const long K = 2 * vol, kc = overlap_bf16x3_kchunk(vol);
  const long nchunks = (K + kc - 1) / kc;
  return (size_t)nchunks * (mr * 16) * ((long)ncfg * 16 * ntt + 64);
}

hipError_t launch_split_rows(const float2* phi, void* out, int ns, long vol, hipStream_t s,
                             int parts) {
  if (parts != 2 && parts != 3) return hipErrorInvalidValue;
  const long n = (long)ns * vol;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_split_rows, dim3((unsigned)blocks), dim3(256), 0, s, phi, (__bf16*)out,
                     n, vol, parts);
  return hipGetLastError();
}

// ------------------------------------------------------------------ overlaps, shared sinks
// The sink-row form with the sinks staged in LDS and shared by the four waves of a workgroup,
// each wave one tile of 16 NT orbital columns over the same K chunk: the split sinks cross
// from L2 once per four tiles instead of once per wave and step, which leaves the orbital
// stream from HBM as the only large flow.  A stage is KS steps of 32 K values of every sink row and
// part; the next stage's sink loads go out before the current stage's products and are
// written to the other LDS buffer after them (one barrier per stage).  LDS layout per part
// and 8-value K group q: R rows of 16 bytes, plane stride R 16 + 16 bytes, so the 16 rows a
// quarter-wave reads are one contiguous 256-byte run (no bank conflicts) and the loaders'
// 16-byte writes from four consecutive rows of 16 K groups land in distinct banks.
// Steps past the chunk end read zero from both operands and add nothing.  Grid (group of four
// column tiles, K chunk) with the group fastest, so that the workgroups resident at one time
// read one chunk of the split sinks, which stays in L2.
// Three forms (FORM): 0, stages of 2 steps with 2 steps of orbital loads in flight, small
// enough for two workgroups per CU (two waves per SIMD hide each other's load latency);
// 1, stages of 4 steps with 4 steps in flight, one workgroup per CU; 2, stages of 2 steps with
// 4 steps in flight, two workgroups per CU.  A shape whose 4-step stages do not fit the LDS
// budget uses 2-step stages.
template <int MR, int TERMS, int FORM>
constexpr int ovl_lds_ks() {
  constexpr int npa = TERMS == 6 ? 3 : 2, r = MR * 16;
  return FORM == 1 && 2 * npa * 16 * (r * 16 + 16) <= 160 * 1024 ? 4 : 2;
}

template <int MR, int NT, int TERMS, int FORM>
__global__ __launch_bounds__(256, FORM == 1 ? 1 : 2) void k_overlap_lds(
    const __bf16* __restrict__ a0p, const __bf16* __restrict__ a1p,
    const __bf16* __restrict__ a2p, const float2* __restrict__ psi, float* __restrict__ partial,
    int nrow, int nt, long ntp, long vol, int ncfg, long kchunk, int nchunks, long cstride) {
  constexpr int NPA = TERMS == 6 ? 3 : 2, R = MR * 16, KS = ovl_lds_ks<MR, TERMS, FORM>();
  constexpr int PD = FORM == 0 ? 2 : 4, Q = KS * 4;
  constexpr int SU = PD > KS ? PD / KS : 1;  // stages per loop iteration (static buffers)
  constexpr int PL = R * 16 + 16;           // bytes per (part, K group) plane
  constexpr int STAGE = NPA * Q * PL;       // bytes per stage
  constexpr int NCH = R * Q;                // 16-byte chunks per part and stage
  constexpr int NLD = (NCH + 255) / 256;    // per thread and part
  static_assert((KS * SU) % PD == 0, "the orbital ring indexes by step within an iteration");
  __shared__ __attribute__((aligned(16))) unsigned char lds[2 * STAGE];
  using u32x4 = __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int row = lane & 15, grp = lane >> 4;
  // the wave's tile: columns col0 ... col0 + 16 NT - 1, column (c, t) = c ntp + t (t < nt)
  const long tile = (long)blockIdx.x * 4 + wave;
  const long ncol = (long)ncfg * ntp, col0 = tile * 16 * NT;
  const bool live = col0 < ncol;
  const long cf = live ? col0 / ntp : 0;  // the tile's first and last configurations
  const long cl = live ? ((col0 + 16 * NT < ncol ? col0 + 16 * NT : ncol) - 1) / ntp : 0;
  const int chunk = blockIdx.y;
  const long K = 2 * vol;
  const long k0 = (long)chunk * kchunk;
  const long k1 = k0 + kchunk < K ? k0 + kchunk : K;
  const int nst = (int)((k1 - k0 + 32 * KS - 1) / (32 * KS));
  const uint32_t abytes = (uint32_t)(nrow * K * 2);
  const rsrc_t ra[3] = {make_rsrc(a0p, abytes), make_rsrc(a1p, abytes),
                        make_rsrc(a2p, TERMS == 6 ? abytes : 0u)};
  const rsrc_t rb = make_rsrc(psi + cf * cstride,
                              live ? (uint32_t)(((cl - cf) * cstride + (long)nt * vol) * 8) : 0u);
  constexpr uint32_t kFar = 0x7FFFFFF0u;  // past any resource: loads return zero
  // the stage loader, part by part (the part selects the buffer resource, so it stays
  // uniform): chunk id = ((row / 4) Q + q) 4 + row % 4
  u32x4 apf[NPA][NLD];
  auto loadA = [&](long ks) {
#pragma unroll
    for (int j = 0; j < NPA; ++j)
#pragma unroll
      for (int i = 0; i < NLD; ++i) {
        const int id = tid + 256 * i;
        if (NCH % 256 != 0 && id >= NCH) break;
        const int rlo = id & 3, q = (id >> 2) % Q, r = (id >> 2) / Q * 4 + rlo;
        const long k = ks + 8 * q;
        const uint32_t o = (r < nrow && k < k1) ? (uint32_t)(r * K + k) * 2u : kFar;
        apf[j][i] = bload<u32x4>(ra[j], o, 0u);
      }
  };
  auto storeA = [&](int buf) {
#pragma unroll
    for (int j = 0; j < NPA; ++j)
#pragma unroll
      for (int i = 0; i < NLD; ++i) {
        const int id = tid + 256 * i;
        if (NCH % 256 != 0 && id >= NCH) break;
        const int rlo = id & 3, q = (id >> 2) % Q, r = (id >> 2) / Q * 4 + rlo;
        *reinterpret_cast<u32x4*>(lds + buf * STAGE + (j * Q + q) * PL + r * 16) = apf[j][i];
      }
  };
  uint32_t ob[NT];
#pragma unroll
  for (int b = 0; b < NT; ++b) {
    const long col = col0 + b * 16 + row, cc = col / ntp, t = col - cc * ntp;
    ob[b] = (live && col < ncol && t < nt)
                ? (uint32_t)(((cc - cf) * cstride + t * vol + 4 * grp) * 8)
                : kFar;
  }
  struct BStep {
    float4 b0[NT], b1[NT];
  };
  auto loadB = [&](BStep& st, long kb) {
    const uint32_t sb = (uint32_t)kb * 4u;
    const bool in = kb + 8 * grp < k1;
#pragma unroll
    for (int b = 0; b < NT; ++b) {
      const uint32_t o = in ? ob[b] : kFar;
      st.b0[b] = bload<float4>(rb, o, sb);
      st.b1[b] = bload<float4>(rb, o + 16u, sb);
    }
  };
  f32x4 acc[MR][NT];
#pragma unroll
  for (int a = 0; a < MR; ++a)
#pragma unroll
    for (int b = 0; b < NT; ++b) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
  // one step: the sink fragments of K group q = 4 u + grp from LDS, the orbital values split
  auto compute = [&](int buf, int u, const BStep& B) {
    const unsigned char* base = lds + buf * STAGE + (4 * u + grp) * PL + row * 16;
    bf16x8 av[NPA][MR];
#pragma unroll
    for (int j = 0; j < NPA; ++j)
#pragma unroll
      for (int a = 0; a < MR; ++a)
        av[j][a] = __builtin_bit_cast(
            bf16x8, *reinterpret_cast<const u32x4*>(base + j * Q * PL + a * 256));
#pragma unroll
    for (int b = 0; b < NT; ++b) {
      const float v[8] = {B.b0[b].x, B.b0[b].y, B.b0[b].z, B.b0[b].w,
                          B.b1[b].x, B.b1[b].y, B.b1[b].z, B.b1[b].w};
      bf16x8 pr[NPA];
      split8_pk<NPA>(v, pr);
      constexpr int TA[6] = {0, 0, 1, 0, 1, 2}, TB[6] = {0, 1, 0, 2, 1, 0};
#pragma unroll
      for (int term = 0; term < TERMS; ++term)
#pragma unroll
        for (int a = 0; a < MR; ++a)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[TA[term]][a], pr[TB[term]],
                                                              acc[a][b], 0, 0, 0);
    }
  };
  BStep B[PD];
  loadA(k0);
#pragma unroll
  for (int j = 0; j < PD; ++j) loadB(B[j], k0 + 32 * j);
  storeA(0);
  __syncthreads();
  for (int st0 = 0; st0 < nst; st0 += SU) {
#pragma unroll
    for (int h = 0; h < SU; ++h) {
      const int st = st0 + h;
      if (st >= nst) break;
      const int buf = SU == 2 ? h : (st & 1);
      const long ks = k0 + (long)st * 32 * KS;
      const bool more = st + 1 < nst;
      if (more) loadA(ks + 32 * KS);
#pragma unroll
      for (int u = 0; u < KS; ++u) {
        const int i = (h * KS + u) % PD;
        compute(buf, u, B[i]);
        loadB(B[i], ks + 32 * (u + PD));
      }
      if (more) storeA(buf ^ 1);
      __syncthreads();
    }
  }
  if (!live) return;
  constexpr int ldp = NT * 16;
  float* out = partial + (tile * nchunks + chunk) * (long)R * ldp;
#pragma unroll
  for (int a = 0; a < MR; ++a)
#pragma unroll
    for (int b = 0; b < NT; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) out[(long)(a * 16 + grp * 4 + r) * ldp + b * 16 + row] = acc[a][b][r];
}

hipError_t launch_overlap_lds(const void* rows_split, const float2* psi, double2* M, int ns,
                              int nt, long vol, int ncfg, float* partial, hipStream_t s,
                              int terms, int form, long cstride) {
  if ((terms != 3 && terms != 6) || form < 0 || form > 2) return hipErrorInvalidValue;
  if (cstride == 0) cstride = nt * vol;
  if (ns < 1 || ns > 16 * OV_MAXT || nt < 1 || nt > 16 * OV_MAXT || (vol & 3))
    return hipErrorInvalidValue;
  const int nrow = 2 * ns, mr = (nrow + 15) / 16;
  const OvlCols cols = ovl_cols(nt, mr, vol, ncfg, cstride);
  const int ntt = cols.nt_tiles;
  const long K = 2 * vol, kc = overlap_bf16x3_kchunk(vol);
  const int nchunks = (int)((K + kc - 1) / kc);
  const __bf16* a0 = (const __bf16*)rows_split;
  const long plane = (long)nrow * K;
  const __bf16* a1 = a0 + plane;
  const __bf16* a2 = a1 + plane;
  const dim3 grid((unsigned)((cols.ntiles + 3) / 4), (unsigned)nchunks), block(256);
  bool launched = false;
#define NRC_OL_T(A, B, T, F)                                                                  \
  hipLaunchKernelGGL((k_overlap_lds<A, B, T, F>), grid, block, 0, s, a0, a1, T == 6 ? a2 : a1,  \
                     psi, partial, nrow, nt, cols.ntp, vol, ncfg, kc, nchunks, cstride)
#define NRC_OL_CASE(A, B)                                                                     \
  if (!launched && mr == A && ntt == B) {                                                     \
    launched = true;                                                                          \
    if (terms == 3 && form == 0) NRC_OL_T(A, B, 3, 0);                                        \
    else if (terms == 3 && form == 1) NRC_OL_T(A, B, 3, 1);                                   \
    else if (terms == 3) NRC_OL_T(A, B, 3, 2);                                                \
    else if (form == 0) NRC_OL_T(A, B, 6, 0);                                                 \
    else if (form == 1) NRC_OL_T(A, B, 6, 1);                                                 \
    else NRC_OL_T(A, B, 6, 2);                                                                \
  }
#define NRC_OL_ROW(A) NRC_OL_CASE(A, 1) NRC_OL_CASE(A, 2) NRC_OL_CASE(A, 3) NRC_OL_CASE(A, 4)
  NRC_OL_ROW(1) NRC_OL_ROW(2) NRC_OL_ROW(3) NRC_OL_ROW(4) NRC_OL_ROW(5) NRC_OL_ROW(6)
  NRC_OL_ROW(7) NRC_OL_ROW(8)
#undef NRC_OL_ROW
#undef NRC_OL_CASE
#undef NRC_OL_T
  if (!launched) return hipErrorInvalidValue;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long n = (long)ncfg * ns * nt;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_overlap_rows_reduce, dim3((unsigned)blocks), dim3(256), 0, s, partial, M,
                     ns, nt, cols.ntp, nchunks, mr, ntt, (long)ncfg);
  return hipGetLastError();
}

size_t pair_bf16x3_partial_floats(int n1, int n2, long vol, int ncfg) {
  const int mt = (n1 + 15) / 16, nt = (n2 + 15) / 16;
  const long K = 2 * vol, kc = overlap_bf16x3_kchunk(vol);
  const long nchunks = (K + kc - 1) / kc;
  return (size_t)ncfg * nchunks * (mt * 16) * (nt * 32);
}

hipError_t launch_pair_bf16x3(const float2* psi1, const float2* psi2, const double* wf,
                              const int* opp, double2* M, int n1, int n2, long vol, int ncfg,
                              float* partial, hipStream_t s) {
  const bool cj = opp == nullptr;
  if (n1 < 1 || n1 > 16 * OV_MAXT || n2 < 1 || n2 > 64 || (vol & 3))
    return hipErrorInvalidValue;
  const int mt = (n1 + 15) / 16, nt = (n2 + 15) / 16;
  const long K = 2 * vol, kc = overlap_bf16x3_kchunk(vol);
  const int nchunks = (int)((K + kc - 1) / kc);
  const dim3 grid(nchunks, ncfg), block(64 * OV_WAVES);
#define NRC_PR_CASE(A, B)                                                                     \
  if (mt == A && nt == B && !cj) {                                                            \
    hipLaunchKernelGGL((k_pair_bf16x3<A, B, false>), grid, block, 0, s, psi1, psi2, wf, opp,   \
                       partial, n1, n2, vol, mt, nt, kc);                                     \
  } else if (mt == A && nt == B) {                                                            \
    hipLaunchKernelGGL((k_pair_bf16x3<A, B, true>), grid, block, 0, s, psi1, psi2, wf, opp,    \
                       partial, n1, n2, vol, mt, nt, kc);                                     \
  }
#define NRC_PR_ROW(A) NRC_PR_CASE(A, 1) NRC_PR_CASE(A, 2) NRC_PR_CASE(A, 3) NRC_PR_CASE(A, 4)
  NRC_PR_ROW(1) NRC_PR_ROW(2) NRC_PR_ROW(3) NRC_PR_ROW(4)
#undef NRC_PR_ROW
#undef NRC_PR_CASE
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long n = (long)ncfg * n1 * n2;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  // column pairs (2t, 2t + 1) over 2 nt tiles of 16: the overlap's reduction applies
  hipLaunchKernelGGL(k_overlap_reduce, dim3((unsigned)blocks), dim3(256), 0, s, partial, M, n1,
                     n2, nchunks, mt, 2 * nt, (long)ncfg);
  return hipGetLastError();
}

// units of the mirror form: L/2 - 1 plane pairs and the planes 0 and L/2
static int pair_gh_units(int L) { return L / 2 + 1; }

size_t pair_gh_bf16x3_partial_floats(int n, long vol, int ncfg, int L) {
  const size_t chunks = 2 * pair_bf16x3_partial_floats(n, n, vol, ncfg);
  if (L <= 0) return chunks;
  const int mt = (n + 15) / 16;
  const size_t units = 2 * (size_t)ncfg * pair_gh_units(L) * (mt * 16) * (mt * 32);
  return std::max(chunks, units);
}

hipError_t launch_pair_gh_bf16x3(const float2* psi, const double* wf, const int* opp,
                                 double2* G, double2* H, int n, long vol, int ncfg,
                                 float* partial, hipStream_t s, int L, const float* ws,
                                 bool ws_neg) {
  if (n < 1 || n > 48 || (vol & 3)) return hipErrorInvalidValue;
  if (L > 0 && ((long)L * L * L != vol || L % 4 || L < 8 || ws == nullptr))
    return hipErrorInvalidValue;
  const int mt = (n + 15) / 16;
  const long K = 2 * vol, kc = overlap_bf16x3_kchunk(vol);
  const int nchunks = L > 0 ? pair_gh_units(L) : (int)((K + kc - 1) / kc);
  const dim3 grid(nchunks, ncfg), block(64 * OV_WAVES);
#define NRC_PGM(A)                                                                            \
  if (ws_neg)                                                                                 \
    hipLaunchKernelGGL((k_pair_gh_mirror<A, true>), grid, block, 0, s, psi, ws, partial, n, L); \
  else                                                                                        \
    hipLaunchKernelGGL((k_pair_gh_mirror<A, false>), grid, block, 0, s, psi, ws, partial, n, L);
  if (L > 0 && mt == 1) {
    NRC_PGM(1)
  } else if (L > 0 && mt == 2) {
    NRC_PGM(2)
  } else if (L > 0) {
    NRC_PGM(3)
  }
#undef NRC_PGM
  else if (mt == 1)
    hipLaunchKernelGGL(k_pair_gh_bf16x3<1>, grid, block, 0, s, psi, wf, opp, partial, n, vol, kc);
  else if (mt == 2)
    hipLaunchKernelGGL(k_pair_gh_bf16x3<2>, grid, block, 0, s, psi, wf, opp, partial, n, vol, kc);
  else
    hipLaunchKernelGGL(k_pair_gh_bf16x3<3>, grid, block, 0, s, psi, wf, opp, partial, n, vol, kc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long nout = (long)ncfg * n * n;
  long blocks = (nout + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  if (L > 0) {
    hipLaunchKernelGGL(k_pair_gh_reduce, dim3((unsigned)blocks), dim3(256), 0, s, partial, G, H,
                       n, nchunks, L / 2 - 1, mt, (long)ncfg);
    return hipGetLastError();
  }
  const size_t half = (size_t)ncfg * nchunks * (mt * 16) * (mt * 32);
  hipLaunchKernelGGL(k_overlap_reduce, dim3((unsigned)blocks), dim3(256), 0, s, partial, G, n, n,
                     nchunks, mt, 2 * mt, (long)ncfg);
  hipLaunchKernelGGL(k_overlap_reduce, dim3((unsigned)blocks), dim3(256), 0, s, partial + half, H,
                     n, n, nchunks, mt, 2 * mt, (long)ncfg);
  return hipGetLastError();
}

}  // namespace nrc
