// proposal.hip -- the learned auxiliary-field proposal: a 3-layer periodic 3D convolutional
// network (3x3x3 kernels, 16 hidden channels, BF16 weights and activations, FP32
// accumulation) that maps the previous slice's field phi(t-1) and the configuration's current
// normalised log-weight u to per-site Bernoulli logits for phi(t), its sampler with exact
// reweighting ratios, and its training kernels (weighted maximum likelihood, Adam).
//
// Lattice sites are the reference's order z + L (y + L x) (src/utils/vector3.C:175-187); a
// convolution is out[s] = sum_tap W[tap] in[s + d(tap)] with periodic wrap, tap =
// 9 (dx+1) + 3 (dy+1) + (dz+1).  Activations are [config][site][16] bf16 (32 B per site).
//
// Parameters (FP32 master copy, P_N floats):
//   W1 [16][27][2] (input 0: phi(t-1) in {-1, 0, +1}, input 1: u), b1 [16]   -> ReLU
//   W2 [16][27][16], b2 [16]                                                   -> ReLU
//   W3 [27][16], b3                                                            -> logit
// The 16 -> 16 layer, forward and backward-data, runs on the matrix cores
// (v_mfma_f32_16x16x32_bf16); the 2 -> 16 and 16 -> 1 layers and the weight gradients are
// VALU kernels.
#include <type_traits>

#include "nrc_device.h"
#include "nrc_internal.h"

namespace nrc {

constexpr int PC = 16;
constexpr int NTAP = 27;
constexpr int P_W1 = 0;
constexpr int P_B1 = P_W1 + PC * NTAP * 2;
constexpr int P_W2 = P_B1 + PC;
constexpr int P_B2 = P_W2 + PC * NTAP * PC;
constexpr int P_W3 = P_B2 + PC;
constexpr int P_B3 = P_W3 + NTAP * PC;
constexpr int P_N = P_B3 + 1;
static_assert(P_N == 8241, "parameter count");

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

struct Lat {
  int L;
  long V;
};

// neighbour of site s at tap (periodic)
__device__ __forceinline__ long nbr_site(const Lat& g, int x, int y, int z, int tap) {
  const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
  int xx = x + dx, yy = y + dy, zz = z + dz;
  xx += (xx < 0) ? g.L : 0;
  xx -= (xx >= g.L) ? g.L : 0;
  yy += (yy < 0) ? g.L : 0;
  yy -= (yy >= g.L) ? g.L : 0;
  zz += (zz < 0) ? g.L : 0;
  zz -= (zz >= g.L) ? g.L : 0;
  return ((long)xx * g.L + yy) * g.L + zz;
}

__device__ __forceinline__ void site_xyz(const Lat& g, long s, int& x, int& y, int& z) {
  z = (int)(s % g.L);
  const long r = s / g.L;
  y = (int)(r % g.L);
  x = (int)(r / g.L);
}

__device__ __forceinline__ float bf(const __bf16 v) { return (float)v; }

// The 27 neighbours of a site as sums of three per-axis terms (wrapped once per site):
// neighbour(tap) = ox[tap / 9] + oy[(tap / 3) % 3] + oz[tap % 3].  With tap a compile-time
// constant this is two additions per neighbour.
struct NbrOff {
  long ox[3];
  int oy[3], oz[3];
  __device__ __forceinline__ void init(const Lat& g, int x, int y, int z) {
    const int xm = x == 0 ? g.L - 1 : x - 1, xp = x == g.L - 1 ? 0 : x + 1;
    const int ym = y == 0 ? g.L - 1 : y - 1, yp = y == g.L - 1 ? 0 : y + 1;
    const int zm = z == 0 ? g.L - 1 : z - 1, zp = z == g.L - 1 ? 0 : z + 1;
    const long LL = (long)g.L * g.L;
    ox[0] = xm * LL; ox[1] = x * LL; ox[2] = xp * LL;
    oy[0] = ym * g.L; oy[1] = y * g.L; oy[2] = yp * g.L;
    oz[0] = zm; oz[1] = z; oz[2] = zp;
  }
  template <int TAP>
  __device__ __forceinline__ long at() const {
    return ox[TAP / 9] + oy[(TAP / 3) % 3] + oz[TAP % 3];
  }
};

template <int TAP, typename F>
__device__ __forceinline__ void for_taps(F&& f) {
  if constexpr (TAP < 27) {
    f(std::integral_constant<int, TAP>{});
    for_taps<TAP + 1>(f);
  }
}

// ------------------------------------------------------------------ parameters
__global__ void k_fill_zero_f(float* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.f;
}
__global__ void k_fill_zero_d(double* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.0;
}

// bf16 copies of the master weights, and W2 flipped and transposed for the backward-data
// convolution: w2t[ci][tap][co] = W2[co][26 - tap][ci]
__global__ void k_refresh(const float* p, __bf16* wb, __bf16* w2t) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P_N; i += gridDim.x * blockDim.x) {
    wb[i] = (__bf16)p[i];
    if (i >= P_W2 && i < P_B2) {
      const int j = i - P_W2, co = j / (NTAP * PC), tap = (j / PC) % NTAP, ci = j % PC;
      w2t[(ci * NTAP + (NTAP - 1 - tap)) * PC + co] = (__bf16)p[i];
    }
  }
}

// SGD with momentum 0.9: m = 0.9 m + g, p -= lr m.  Unlike Adam, parameters whose gradient is
// pure noise move only in proportion to it.
__global__ void k_sgdm(float* p, const float* g, float* m, int n, float lr) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float mi = 0.9f * m[i] + g[i];
    m[i] = mi;
    p[i] -= lr * mi;
  }
}

// Adam (Kingma & Ba 2015) with bias correction; step is 1-based
__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1,
                       float b2, float eps, int step) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// ------------------------------------------------------------------ layer 1 (2 -> 16)
// The phi(t-1) channel is a GEMM with M = 16 output channels, N = 16 sites and K = 27 taps
// (padded to 32): one v_mfma_f32_16x16x32_bf16 per tile, lane l supplying the weights of
// channel l % 16 and the eight neighbours 8 (l / 16) .. +7 of site l % 16 as +-1 (0: no
// previous slice, or a padded tap).  The u channel is constant over the lattice, so its
// contribution is u * sum_tap W1[co][tap][1]; bias, ReLU, bf16.
// prev: [B][nw] field words of slice t-1 (null: no previous slice); u: [B]
__global__ __launch_bounds__(256) void k_conv1(const uint32_t* __restrict__ prev, int nw,
                                               const float* __restrict__ u,
                                               const __bf16* __restrict__ wb,
                                               __bf16* __restrict__ A1, Lat g, long B) {
  const int lane = threadIdx.x & 63;
  const int row = lane & 15, grp = lane >> 4;
  bf16x8 wa;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int tap = 8 * grp + k;
    wa[k] = tap < NTAP ? wb[P_W1 + (row * NTAP + tap) * 2] : (__bf16)0.f;
  }
  float bco[4], sco[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int co = grp * 4 + r;
    float su = 0.f;
    for (int tap = 0; tap < NTAP; ++tap) su += bf(wb[P_W1 + (co * NTAP + tap) * 2 + 1]);
    sco[r] = su;
    bco[r] = bf(wb[P_B1 + co]);
  }
  const long ntiles = B * g.V / 16;
  const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long nwave = (long)gridDim.x * (blockDim.x >> 6);
  for (long tile = wave; tile < ntiles; tile += nwave) {
    const long i0 = tile * 16;
    const long c = i0 / g.V;
    const long s = i0 - c * g.V + row;
    bf16x8 b;
    if (prev) {
      int x, y, z;
      site_xyz(g, s, x, y, z);
      NbrOff nb;
      nb.init(g, x, y, z);
      const uint32_t* w = prev + c * nw;
      // taps 8 grp + k: the group picks one of four compile-time sets of eight taps
      auto fill = [&](auto base_tag) {
        constexpr int T0 = decltype(base_tag)::value;
#pragma unroll
        for (int k = 0; k < 8; ++k) b[k] = (__bf16)0.f;
        for_taps<0>([&](auto tap_tag) {
          constexpr int TAP = decltype(tap_tag)::value;
          if constexpr (TAP >= T0 && TAP < T0 + 8) {
            const long q = nb.at<TAP>();
            b[TAP - T0] = ((w[q >> 5] >> (q & 31)) & 1u) ? (__bf16)1.f : (__bf16)-1.f;
          }
        });
      };
      if (grp == 0) fill(std::integral_constant<int, 0>{});
      else if (grp == 1) fill(std::integral_constant<int, 8>{});
      else if (grp == 2) fill(std::integral_constant<int, 16>{});
      else fill(std::integral_constant<int, 24>{});
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) b[k] = (__bf16)0.f;
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, b, acc, 0, 0, 0);
    const float uc = u[c];
    typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
    bf16x4 res;
#pragma unroll
    for (int r = 0; r < 4; ++r) res[r] = (__bf16)fmaxf(acc[r] + bco[r] + uc * sco[r], 0.f);
    *reinterpret_cast<bf16x4*>(A1 + (c * g.V + s) * PC + grp * 4) = res;
  }
}

// ------------------------------------------------------------------ 16 -> 16 on the matrix cores
// out[c][s][co] = sum_tap sum_ci w[co][tap][ci] in[c][s + d(tap)][ci] as a GEMM with M = 16
// output channels, N = 16 sites per tile and K = 27 taps x 16 channels (14 steps of 32, the
// last half zero).  v_mfma_f32_16x16x32_bf16: lane l supplies A row l % 16 and B column
// l % 16 with the eight K values 8 (l / 16) .. +7 of the step, i.e. tap 2 step + (l / 16) / 2
// and channels 8 ((l / 16) % 2) .. +7; the weights of a lane stay in registers.
// FWD: out = bf16(relu(acc + bias)); otherwise out = bf16(acc) where mask > 0, else 0.
template <bool FWD>
__global__ __launch_bounds__(256) void k_conv16(const __bf16* __restrict__ in,
                                                const __bf16* __restrict__ w,
                                                const __bf16* __restrict__ bias,
                                                const __bf16* __restrict__ mask,
                                                __bf16* __restrict__ out, Lat g, long B) {
  const int lane = threadIdx.x & 63;
  const int row = lane & 15, grp = lane >> 4;
  const int ci0 = 8 * (grp & 1);
  bf16x8 wa[14];
#pragma unroll
  for (int st = 0; st < 14; ++st) {
    const int tap = 2 * st + (grp >> 1);
    if (tap < NTAP) {
      wa[st] = *reinterpret_cast<const bf16x8*>(w + (row * NTAP + tap) * PC + ci0);
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) wa[st][k] = (__bf16)0.f;
    }
  }
  float bco[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) bco[r] = FWD ? bf(bias[grp * 4 + r]) : 0.f;
  const long ntiles = B * g.V / 16;
  const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long nwave = (long)gridDim.x * (blockDim.x >> 6);
  for (long tile = wave; tile < ntiles; tile += nwave) {
    const long i0 = tile * 16;
    const long c = i0 / g.V;
    const long s = i0 - c * g.V + row;
    int x, y, z;
    site_xyz(g, s, x, y, z);
    NbrOff nb;
    nb.init(g, x, y, z);
    const __bf16* inc = in + c * g.V * PC + ci0;
    const bool odd = (grp >> 1) != 0;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    auto step = [&](auto st_tag) {
      constexpr int ST = decltype(st_tag)::value;
      bf16x8 b;
      if constexpr (2 * ST + 1 < NTAP) {
        const long q = odd ? nb.at<2 * ST + 1>() : nb.at<2 * ST>();
        b = *reinterpret_cast<const bf16x8*>(inc + q * PC);
      } else {  // tap 26 and the zero padding tap 27
        if (!odd) {
          b = *reinterpret_cast<const bf16x8*>(inc + nb.at<2 * ST>() * PC);
        } else {
#pragma unroll
          for (int k = 0; k < 8; ++k) b[k] = (__bf16)0.f;
        }
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[ST], b, acc, 0, 0, 0);
    };
    step(std::integral_constant<int, 0>{});
    step(std::integral_constant<int, 1>{});
    step(std::integral_constant<int, 2>{});
    step(std::integral_constant<int, 3>{});
    step(std::integral_constant<int, 4>{});
    step(std::integral_constant<int, 5>{});
    step(std::integral_constant<int, 6>{});
    step(std::integral_constant<int, 7>{});
    step(std::integral_constant<int, 8>{});
    step(std::integral_constant<int, 9>{});
    step(std::integral_constant<int, 10>{});
    step(std::integral_constant<int, 11>{});
    step(std::integral_constant<int, 12>{});
    step(std::integral_constant<int, 13>{});
    // D: lane holds output channels 4 grp .. 4 grp + 3 of site s
    const long o = (c * g.V + s) * PC + grp * 4;
    typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
    bf16x4 res;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float v = acc[r];
      if constexpr (FWD) {
        v = fmaxf(v + bco[r], 0.f);
      } else {
        v = bf(mask[o + r]) > 0.f ? v : 0.f;
      }
      res[r] = (__bf16)v;
    }
    *reinterpret_cast<bf16x4*>(out + o) = res;
  }
}

// ------------------------------------------------------------------ layer 3 (16 -> 1)
typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2;

// logit of a site: b3 + sum_tap sum_ci W3[tap][ci] A2[s + d(tap)][ci], bf16 pairs through
// v_dot2 with FP32 accumulation; w3 holds the weights as 216 bf16 pairs
__device__ __forceinline__ float conv3_logit(const __bf16* __restrict__ A2c, const NbrOff& nb,
                                             const bf16x2* w3, float b3) {
  float acc = b3;
  for_taps<0>([&](auto tap_tag) {
    constexpr int TAP = decltype(tap_tag)::value;
    const bf16x2* p = reinterpret_cast<const bf16x2*>(A2c + nb.at<TAP>() * PC);
    typedef __attribute__((__vector_size__(8 * sizeof(uint32_t)))) uint32_t u32x8;
    const u32x8 raw = *reinterpret_cast<const u32x8*>(p);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t word = raw[k];  // a scalar copy: bit_cast of a vector element reads element 0
      acc = __builtin_amdgcn_fdot2_f32_bf16(w3[TAP * 8 + k], __builtin_bit_cast(bf16x2, word),
                                            acc, false);
    }
    // groups of nine taps: the loads of later groups are not hoisted over this one
    if constexpr (TAP % 9 == 8) __builtin_amdgcn_sched_barrier(0);
  });
  return acc;
}

// sum over a wavefront in FP64
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// MODE 0: logits out.  MODE 1: sample phi(t) from Philox(seed; site / 4, t, config, stream)
// word site % 4: phi = +1 iff the word is below T = proposal_threshold(logit) (about
// sigmoid(logit) 2^32), so q(+1) = T / 2^32 exactly; the log ratio log(q_uniform / q) of every
// site is summed in FP64 into logr[c].  MODE 2: training, dlogit = scale w[c] (q(+1) - phi01) for
// the target field and the weighted negative log-likelihood summed into loss[0].
// Sites of a wavefront are consecutive and of one configuration (V % 64 == 0).
template <int MODE>
__global__ __launch_bounds__(256) void k_conv3(const __bf16* __restrict__ A2,
                                               const __bf16* __restrict__ wb, Lat g, long B,
                                               float* __restrict__ logits, uint64_t seed,
                                               uint32_t config0, uint32_t t, uint32_t stream,
                                               uint32_t* __restrict__ words,
                                               double* __restrict__ logr,
                                               const uint32_t* __restrict__ target,
                                               const float* __restrict__ w, float scale,
                                               double* __restrict__ loss) {
  __shared__ bf16x2 w3[NTAP * PC / 2];
  for (int i = threadIdx.x; i < NTAP * PC / 2; i += blockDim.x) {
    bf16x2 pr;
    pr[0] = wb[P_W3 + 2 * i];
    pr[1] = wb[P_W3 + 2 * i + 1];
    w3[i] = pr;
  }
  __syncthreads();
  const float b3 = bf(wb[P_B3]);
  const long n = B * g.V;
  const int nw = (int)(g.V / 32);
  // every thread of a wavefront runs the same number of iterations (n % 64 == 0)
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long c = i / g.V, s = i - c * g.V;
    int x, y, z;
    site_xyz(g, s, x, y, z);
    NbrOff nb;
    nb.init(g, x, y, z);
    const float lg = conv3_logit(A2 + c * g.V * PC, nb, w3, b3);
    if constexpr (MODE == 0) {
      logits[i] = lg;
    } else if constexpr (MODE == 1) {
      uint32_t ctr[4] = {(uint32_t)(s >> 2), t, config0 + (uint32_t)c, stream};
      philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
      const uint32_t r = ctr[s & 3];
      const uint32_t th = proposal_threshold(lg);
      const bool up = r < th;
      // log(q_uniform / q(phi)) = log(2^31 / T) or log(2^31 / (2^32 - T))
      RatioAcc ra;
      ra.add(up ? th : (uint32_t)(0u - th));
      const double lr = ra.value();
      const uint64_t m = __ballot(up);
      const int lane = threadIdx.x & 63;
      if (lane == 0) words[c * nw + (s >> 5)] = (uint32_t)m;
      if (lane == 32) words[c * nw + (s >> 5)] = (uint32_t)(m >> 32);
      const double tot = wave_sum(lr);
      if (lane == 0) atomicAdd(logr + c, tot);
    } else {
      const uint32_t wd = target[c * nw + (s >> 5)];
      const float y1 = ((wd >> (s & 31)) & 1u) ? 1.f : 0.f;
      const float q = 1.f / (1.f + expf(-lg));
      logits[i] = scale * w[c] * (q - y1);
      // -log q(target) = softplus(-lg) for phi = +1, softplus(lg) for phi = -1
      const float a = y1 > 0.f ? -lg : lg;
      const double nll = (double)(fmaxf(a, 0.f) + log1pf(expf(-fabsf(a))));
      const double tot = wave_sum((double)w[c] * nll);
      if ((threadIdx.x & 63) == 0) atomicAdd(loss, tot);
    }
  }
}

// ------------------------------------------------------------------ fused inference
// The whole network for one configuration inside one workgroup, no activation leaving the CU.
// The lattice is swept plane by plane along x; the 2 -> 16 activations of three consecutive
// planes sit in an LDS ring.  Iteration i (0 .. L + 3) computes layer 1 on plane i - 2, layer 2
// on plane i - 3 and completes the logits of plane i - 4 (planes mod L); the sweep is unrolled
// by three so that ring slots are compile-time offsets.
// Layer 1 reads phi(t-1) from a copy with a periodic halo of one site (one byte per site, the
// high byte of the bf16 +-1), so every tap is the lane's site plus a constant.
// The 16 -> 1 layer is split by the x offset of its taps: for the layer-2 plane p,
//   P[dyz][s][dx] = sum_ci W3[9 dx + dyz][ci] A2[p][s][ci]     (dyz = 3 (dy + 1) + dz + 1)
// comes out of the matrix cores (v_mfma_f32_16x16x16_bf16, whose B operand is exactly the
// layer-2 accumulator layout; A rows ordered (dyz, dx), three rows used of every four), is
// staged in LDS for that plane, and every site gathers R_dx(p)[s] = sum_dyz P[dyz][s +
// d(dyz)][dx].  logit(x) = R_-1(x - 1) + R_0(x) + R_+1(x + 1) + b3 accumulates in registers,
// in that order.  Layers 1 and 2 round to bf16 exactly as k_conv1 / k_conv16 do.
// MODE 0: logits out; MODE 1: sample as k_conv3<1> (same Philox words, thresholds and ratio).
constexpr int FUSED_NT = 512;

template <int L>
struct FusedLds {
  static constexpr int L2 = L * L, V = L2 * L, NW = V / 32, LP = L + 2;
  uint32_t in[NW];
  uint32_t out[NW];
  bf16x8 a1[3][2][L2];  // ring slot, channel half, in-plane site
  f32x4 p[9][L2];
  double red[FUSED_NT / 64];
  uint8_t phi[LP * LP * LP];
};

template <int L, int MODE>
__global__ __launch_bounds__(FUSED_NT) void k_infer(const uint32_t* __restrict__ prev,
                                                    const float* __restrict__ u,
                                                    const __bf16* __restrict__ wb, int B,
                                                    float* __restrict__ logits, uint64_t seed,
                                                    uint32_t config0, uint32_t t,
                                                    uint32_t stream, uint32_t* __restrict__ words,
                                                    double* __restrict__ logr) {
  constexpr int L2 = L * L, V = L2 * L, NW = V / 32, TP = L2 / 16;
  constexpr int LP = L + 2, LP2 = LP * LP, LP3 = LP2 * LP;
  constexpr int NWAVE = FUSED_NT / 64;
  constexpr int TPW = (TP + NWAVE - 1) / NWAVE;    // layer tiles per wavefront
  constexpr int SPT = (L2 + FUSED_NT - 1) / FUSED_NT;  // gathered sites per thread
  constexpr int SLOT = 2 * L2 * 16, HALF = L2 * 16;   // bytes of a ring slot, a channel half
  static_assert(L % 4 == 0, "tiles of 16 in-plane sites and 32-site field words");
  static_assert(sizeof(FusedLds<L>) <= 160 * 1024, "LDS of one CU");
  typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
  __shared__ FusedLds<L> sm;
  char* const a1b = reinterpret_cast<char*>(&sm.a1[0][0][0]);

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int row = lane & 15, grp = lane >> 4;
  const int half = grp & 1;
  const bool odd = (grp >> 1) != 0;

  // layer 1: A row = channel `row`, K = taps 8 grp .. +7 of the phi(t-1) channel
  bf16x8 wa1;
  int d1[8];  // the taps as offsets in the haloed field (taps past 26: the site itself)
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int tap = 8 * grp + k;
    wa1[k] = tap < NTAP ? wb[P_W1 + (row * NTAP + tap) * 2] : (__bf16)0.f;
    d1[k] = tap < NTAP ? (tap / 9 - 1) * LP2 + ((tap / 3) % 3 - 1) * LP + (tap % 3 - 1) : 0;
  }
  float b1c[4], s1c[4], b2c[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int co = grp * 4 + r;
    float su = 0.f;
    for (int tap = 0; tap < NTAP; ++tap) su += bf(wb[P_W1 + (co * NTAP + tap) * 2 + 1]);
    s1c[r] = su;
    b1c[r] = bf(wb[P_B1 + co]);
    b2c[r] = bf(wb[P_B2 + co]);
  }
  // layer 2: as k_conv16
  bf16x8 wa2[14];
#pragma unroll
  for (int st = 0; st < 14; ++st) {
    const int tap = 2 * st + (grp >> 1);
    if (tap < NTAP) {
      wa2[st] = *reinterpret_cast<const bf16x8*>(wb + P_W2 + (row * NTAP + tap) * PC + 8 * half);
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) wa2[st][k] = (__bf16)0.f;
    }
  }
  // layer 3 rows: MFMA m, row `row` = 4 j + dx (dx = 3 unused), dyz = 4 m + j; K = channels
  // 4 grp .. +3
  bf16x4 wa3[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int dyz = 4 * m + (row >> 2), dxi = row & 3;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      wa3[m][k] = (dxi < 3 && dyz < 9) ? wb[P_W3 + (9 * dxi + dyz) * PC + 4 * grp + k]
                                       : (__bf16)0.f;
  }
  const float b3 = bf(wb[P_B3]);

  for (int c = blockIdx.x; c < B; c += gridDim.x) {
    if (prev)
      for (int w = tid; w < NW; w += FUSED_NT) sm.in[w] = prev[(long)c * NW + w];
    if constexpr (MODE == 1)
      for (int w = tid; w < NW; w += FUSED_NT) sm.out[w] = 0u;
    __syncthreads();
    if (prev) {
      for (int h = tid; h < LP3; h += FUSED_NT) {
        const int xp = h / LP2, yp = (h / LP) % LP, zp = h % LP;
        const int x = xp == 0 ? L - 1 : (xp == LP - 1 ? 0 : xp - 1);
        const int y = yp == 0 ? L - 1 : (yp == LP - 1 ? 0 : yp - 1);
        const int z = zp == 0 ? L - 1 : (zp == LP - 1 ? 0 : zp - 1);
        const int q = (x * L + y) * L + z;
        sm.phi[h] = ((sm.in[q >> 5] >> (q & 31)) & 1u) ? 0x3f : 0xbf;
      }
    }
    const float uc = u[c];
    // logits of planes i - 3 and i - 2 in progress, of in-plane sites gsite(j): threads in
    // reverse order from the second site on, so that the wavefronts with fewer layer tiles
    // (the last ones) gather the remainder
    float acc_a[SPT], acc_b[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) acc_a[j] = acc_b[j] = 0.f;
    RatioAcc ratio;
    __syncthreads();

    auto iteration = [&](const int i, auto ph_tag) {
      constexpr int PH = decltype(ph_tag)::value;  // i % 3
      // ---- layer 1 on plane i - 2 into ring slot i % 3
      {
        const int pa = (i - 2 + L) % L;
#pragma unroll 1
        for (int tt = 0; tt < TPW; ++tt) {
          const int tile = wave + tt * NWAVE;
          if (TP % NWAVE != 0 && tile >= TP) break;
          const int yz = tile * 16 + row;
          bf16x8 b;
          if (prev) {
            const int y = yz / L, z = yz % L;
            const int pb = (pa + 1) * LP2 + (y + 1) * LP + (z + 1);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
              const uint16_t hb = (uint16_t)((sm.phi[pb + d1[k]] << 8) | 0x80);
              b[k] = __builtin_bit_cast(__bf16, hb);
            }
          } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) b[k] = (__bf16)0.f;
          }
          f32x4 acc = {0.f, 0.f, 0.f, 0.f};
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa1, b, acc, 0, 0, 0);
          bf16x4 res;
#pragma unroll
          for (int r = 0; r < 4; ++r) res[r] = (__bf16)fmaxf(acc[r] + b1c[r] + uc * s1c[r], 0.f);
          *reinterpret_cast<bf16x4*>(a1b + PH * SLOT + (grp >> 1) * HALF + yz * 16 +
                                     (grp & 1) * 8) = res;
        }
      }
      __syncthreads();
      // ---- layer 2 on plane i - 3 (A1 planes i - 4, i - 3, i - 2 in slots i + 1, i + 2, i)
      if (i >= 2) {
        constexpr int SL[3] = {((PH + 1) % 3) * SLOT, ((PH + 2) % 3) * SLOT, PH * SLOT};
#pragma unroll 1
        for (int tt = 0; tt < TPW; ++tt) {
          const int tile = wave + tt * NWAVE;
          if (TP % NWAVE != 0 && tile >= TP) break;
          const int yz = tile * 16 + row;
          const int y = yz / L, z = yz % L;
          const int oy[3] = {((y + L - 1) % L) * L * 16, y * L * 16, ((y + 1) % L) * L * 16};
          const int hb = half * HALF;
          const int oz[3] = {((z + L - 1) % L) * 16 + hb, z * 16 + hb, ((z + 1) % L) * 16 + hb};
          f32x4 acc = {0.f, 0.f, 0.f, 0.f};
          auto step = [&](auto st_tag) {
            constexpr int ST = decltype(st_tag)::value;
            constexpr int TE = 2 * ST, TO = 2 * ST + 1;
            const int qe = oy[(TE / 3) % 3] + oz[TE % 3];
            bf16x8 b;
            if constexpr (TO < NTAP) {
              const int qo = oy[(TO / 3) % 3] + oz[TO % 3];
              if constexpr (SL[TE / 9] == SL[TO / 9]) {
                b = *reinterpret_cast<const bf16x8*>(a1b + SL[TE / 9] + (odd ? qo : qe));
              } else {
                b = *reinterpret_cast<const bf16x8*>(a1b + (odd ? qo + SL[TO / 9]
                                                                 : qe + SL[TE / 9]));
              }
            } else {  // tap 26 and the zero padding tap 27
              b = *reinterpret_cast<const bf16x8*>(a1b + SL[TE / 9] + qe);
              if (odd) {
#pragma unroll
                for (int k = 0; k < 8; ++k) b[k] = (__bf16)0.f;
              }
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa2[ST], b, acc, 0, 0, 0);
          };
          step(std::integral_constant<int, 0>{});
          step(std::integral_constant<int, 1>{});
          step(std::integral_constant<int, 2>{});
          step(std::integral_constant<int, 3>{});
          step(std::integral_constant<int, 4>{});
          step(std::integral_constant<int, 5>{});
          step(std::integral_constant<int, 6>{});
          step(std::integral_constant<int, 7>{});
          step(std::integral_constant<int, 8>{});
          step(std::integral_constant<int, 9>{});
          step(std::integral_constant<int, 10>{});
          step(std::integral_constant<int, 11>{});
          step(std::integral_constant<int, 12>{});
          step(std::integral_constant<int, 13>{});
          bf16x4 a2;
#pragma unroll
          for (int r = 0; r < 4; ++r) a2[r] = (__bf16)fmaxf(acc[r] + b2c[r], 0.f);
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            f32x4 pm = {0.f, 0.f, 0.f, 0.f};
            pm = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa3[m], a2, pm, 0, 0, 0);
            const int dyz = 4 * m + grp;
            if (dyz < 9) sm.p[dyz][yz] = pm;
          }
        }
      }
      __syncthreads();
      // ---- gather the layer-3 partial sums of plane i - 3; logits of plane i - 4 complete
      if (i >= 2) {
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
          const int yz = j == 0 ? tid : j * FUSED_NT + (FUSED_NT - 1 - tid);
          if (SPT * FUSED_NT != L2 && yz >= L2) continue;
          const int y = yz / L, z = yz % L;
          const int oy[3] = {((y + L - 1) % L) * L, y * L, ((y + 1) % L) * L};
          const int oz[3] = {(z + L - 1) % L, z, (z + 1) % L};
          float r0 = 0.f, r1 = 0.f, r2 = 0.f;
#pragma unroll
          for (int d = 0; d < 9; ++d) {
            const f32x4 v = sm.p[d][oy[d / 3] + oz[d % 3]];
            r0 += v[0];
            r1 += v[1];
            r2 += v[2];
          }
          const float done = acc_a[j] + r2;
          acc_a[j] = acc_b[j] + r1;
          acc_b[j] = r0;
          if (i >= 4) {
            const int s = (i - 4) * L2 + yz;
            const float lg = done + b3;
            if constexpr (MODE == 0) {
              logits[(long)c * V + s] = lg;
            } else {
              uint32_t ctr[4] = {(uint32_t)(s >> 2), t, config0 + (uint32_t)c, stream};
              philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
              const uint32_t r = ctr[s & 3];
              const uint32_t th = proposal_threshold(lg);
              const bool up = r < th;
              ratio.add(up ? th : (uint32_t)(0u - th));
              if (up) atomicOr(&sm.out[s >> 5], 1u << (s & 31));
            }
          }
        }
      }
    };
    for (int i0 = 0; i0 < L + 4; i0 += 3) {
      iteration(i0, std::integral_constant<int, 0>{});
      if (i0 + 1 < L + 4) iteration(i0 + 1, std::integral_constant<int, 1>{});
      if (i0 + 2 < L + 4) iteration(i0 + 2, std::integral_constant<int, 2>{});
    }
    if constexpr (MODE == 1) {
      const double ws = wave_sum(ratio.value());
      if (lane == 0) sm.red[wave] = ws;
    }
    __syncthreads();
    if constexpr (MODE == 1) {
      for (int w = tid; w < NW; w += FUSED_NT) words[(long)c * NW + w] = sm.out[w];
      if (tid == 0) {
        double tot = 0.0;
#pragma unroll
        for (int k = 0; k < NWAVE; ++k) tot += sm.red[k];
        logr[c] += tot;
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ backward
// dZ2[c][s][ci] = (A2 > 0) * sum_tap W3[tap][ci] dlogit[c][s - d(tap)]
__global__ __launch_bounds__(256) void k_conv3_bwd(const float* __restrict__ dlogit,
                                                   const __bf16* __restrict__ A2,
                                                   const __bf16* __restrict__ wb,
                                                   __bf16* __restrict__ dZ2, Lat g, long B) {
  __shared__ float w3[NTAP * PC];
  for (int i = threadIdx.x; i < NTAP * PC; i += blockDim.x) w3[i] = bf(wb[P_W3 + i]);
  __syncthreads();
  const long n = B * g.V;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long c = i / g.V, s = i - c * g.V;
    int x, y, z;
    site_xyz(g, s, x, y, z);
    float acc[PC];
#pragma unroll
    for (int k = 0; k < PC; ++k) acc[k] = 0.f;
    for (int tap = 0; tap < NTAP; ++tap) {
      const long q = nbr_site(g, x, y, z, NTAP - 1 - tap);  // s - d(tap)
      const float d = dlogit[c * g.V + q];
#pragma unroll
      for (int k = 0; k < PC; ++k) acc[k] = fmaf(w3[tap * PC + k], d, acc[k]);
    }
    const bf16x8* a = reinterpret_cast<const bf16x8*>(A2 + i * PC);
    const bf16x8 a0 = a[0], a1 = a[1];
    bf16x8 o0, o1;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      o0[k] = (__bf16)(bf(a0[k]) > 0.f ? acc[k] : 0.f);
      o1[k] = (__bf16)(bf(a1[k]) > 0.f ? acc[8 + k] : 0.f);
    }
    bf16x8* dst = reinterpret_cast<bf16x8*>(dZ2 + i * PC);
    dst[0] = o0;
    dst[1] = o1;
  }
}

// Weight gradients: every workgroup sums its share of the (config, site) range into one row
// of `partial`, a second kernel adds the rows.  Sites are staged in chunks of 64 together
// with their 27 neighbour indices (LDS), so the threads, one per gradient entry, only load
// the operand values.
constexpr int WG_CHUNK = 64;

// LAYER 3: entries (tap, ci) = 432 and the bias: sum dlogit[s] A2[s + d(tap)][ci]
// LAYER 2: entries (co, tap, ci) = 6912 handled as 256 threads (co, ci) x 27 taps, bias by
//          the threads with ci == 0: sum dZ2[s][co] A1[s + d(tap)][ci]
// LAYER 1: entries (co, tap) = 432 for input 0, plus per co the u-weighted sum (input 1,
//          the same for all taps) and the bias: sum dZ1[s][co] phi(s + d(tap)), u dZ1, dZ1
template <int LAYER>
__global__ __launch_bounds__(LAYER == 2 ? 256 : 448) void k_wgrad(
    const float* __restrict__ dlogit, const __bf16* __restrict__ dZ, const __bf16* __restrict__ act,
    const uint32_t* __restrict__ prev, int nw, const float* __restrict__ u, Lat g, long B,
    long per_block, float* __restrict__ partial, int row_len) {
  __shared__ long nb[WG_CHUNK][NTAP];
  __shared__ float dv[WG_CHUNK][PC];
  __shared__ long cfg_of[WG_CHUNK];
  const int tid = threadIdx.x;
  const long n = B * g.V;
  const long i0 = (long)blockIdx.x * per_block;
  const long i1 = i0 + per_block < n ? i0 + per_block : n;
  float acc[NTAP];
#pragma unroll
  for (int k = 0; k < NTAP; ++k) acc[k] = 0.f;
  float accb = 0.f, accu = 0.f;
  for (long base = i0; base < i1; base += WG_CHUNK) {
    const int cnt = (int)(i1 - base < WG_CHUNK ? i1 - base : WG_CHUNK);
    __syncthreads();
    for (int j = tid; j < cnt * NTAP; j += blockDim.x) {
      const int k = j / NTAP, tap = j - k * NTAP;
      const long i = base + k, c = i / g.V, s = i - c * g.V;
      int x, y, z;
      site_xyz(g, s, x, y, z);
      nb[k][tap] = c * g.V + nbr_site(g, x, y, z, tap);
      if (tap == 0) cfg_of[k] = c;
    }
    for (int j = tid; j < cnt * PC; j += blockDim.x) {
      const int k = j / PC, ch = j - k * PC;
      const long i = base + k;
      if constexpr (LAYER == 3) {
        dv[k][ch] = ch == 0 ? dlogit[i] : 0.f;
      } else {
        dv[k][ch] = bf(dZ[i * PC + ch]);
      }
    }
    __syncthreads();
    if constexpr (LAYER == 3) {
      if (tid < NTAP * PC) {
        const int tap = tid / PC, ci = tid - (tid / PC) * PC;
        for (int k = 0; k < cnt; ++k) acc[0] = fmaf(dv[k][0], bf(act[nb[k][tap] * PC + ci]), acc[0]);
      } else if (tid == NTAP * PC) {
        for (int k = 0; k < cnt; ++k) accb += dv[k][0];
      }
    } else if constexpr (LAYER == 2) {
      const int co = tid / PC, ci = tid - (tid / PC) * PC;
      for (int k = 0; k < cnt; ++k) {
        const float d = dv[k][co];
#pragma unroll
        for (int tap = 0; tap < NTAP; ++tap)
          acc[tap] = fmaf(d, bf(act[nb[k][tap] * PC + ci]), acc[tap]);
        accb += d;
      }
    } else {
      if (tid < PC * NTAP) {
        const int co = tid / NTAP, tap = tid - (tid / NTAP) * NTAP;
        if (prev) {
          for (int k = 0; k < cnt; ++k) {
            const long q = nb[k][tap], c = cfg_of[k], sq = q - c * g.V;
            const float f = ((prev[c * nw + (sq >> 5)] >> (sq & 31)) & 1u) ? 1.f : -1.f;
            acc[0] = fmaf(dv[k][co], f, acc[0]);
          }
        }
      } else if (tid < PC * NTAP + PC) {
        const int co = tid - PC * NTAP;
        for (int k = 0; k < cnt; ++k) {
          accb += dv[k][co];
          accu = fmaf(dv[k][co], u[cfg_of[k]], accu);
        }
      }
    }
  }
  float* out = partial + (long)blockIdx.x * row_len;
  if constexpr (LAYER == 3) {
    if (tid < NTAP * PC) out[tid] = acc[0];
    else if (tid == NTAP * PC) out[NTAP * PC] = accb;
  } else if constexpr (LAYER == 2) {
    const int co = tid / PC, ci = tid - (tid / PC) * PC;
#pragma unroll
    for (int tap = 0; tap < NTAP; ++tap) out[(co * NTAP + tap) * PC + ci] = acc[tap];
    if (ci == 0) out[NTAP * PC * PC + co] = accb;
  } else {
    if (tid < PC * NTAP) {
      const int co = tid / NTAP, tap = tid - (tid / NTAP) * NTAP;
      out[(co * NTAP + tap) * 2] = acc[0];
    } else if (tid < PC * NTAP + PC) {
      const int co = tid - PC * NTAP;
      for (int tap = 0; tap < NTAP; ++tap) out[(co * NTAP + tap) * 2 + 1] = accu;
      out[PC * NTAP * 2 + co] = accb;
    }
  }
}

// grad[off + j] = sum over rows of partial[row][j]
__global__ void k_sum_rows(const float* partial, int rows, int row_len, float* grad) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < row_len; j += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int r = 0; r < rows; ++r) s += (double)partial[(long)r * row_len + j];
    grad[j] = (float)s;
  }
}

// u[c] = (2 logabs[c * stride] - mu) / sd
__global__ void k_norm_logw(const double* logabs, int stride, int B, double mu, double sd,
                            float* u) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < B; c += gridDim.x * blockDim.x)
    u[c] = (float)((2.0 * logabs[(long)c * stride] - mu) / sd);
}

// field words of (config[i], t[i]) pairs, the plain sampler's Philox fields
__global__ void k_z2_bits_list(uint64_t seed, const uint32_t* cfg, const uint32_t* ts, int n,
                               uint32_t stream, int nw, uint32_t* words) {
  const int ng = (nw + 3) / 4;
  const long tot = (long)n * ng;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (long)gridDim.x * blockDim.x) {
    const int k = (int)(i / ng), gq = (int)(i - (long)k * ng);
    uint32_t ctr[4] = {(uint32_t)gq, ts[k], cfg[k], stream};
    philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t* out = words + (long)k * nw + 4 * gq;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (4 * gq + q < nw) out[q] = ctr[q];
  }
}

// sums[t][c] = sum over sites of phi(x, t) of the plain sampler's field of configuration
// config0 + c (2 popcount - vol): one workgroup of 256 per (c, t), the words regenerated
__global__ __launch_bounds__(256) void k_z2_sums(uint64_t seed, uint32_t config0, int nconf,
                                                 uint32_t stream, int vol, int* sums) {
  __shared__ int red[4];
  const int c = blockIdx.x;
  const uint32_t t = blockIdx.y;
  const int nw = (vol + 31) / 32, ng = (nw + 3) / 4;
  int pop = 0;
  for (int g = threadIdx.x; g < ng; g += 256) {
    uint32_t ctr[4] = {(uint32_t)g, t, config0 + (uint32_t)c, stream};
    philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int w = 4 * g + q;
      if (w >= nw) break;
      const int valid = vol - 32 * w;  // sites of this word inside the volume
      const uint32_t m = valid >= 32 ? 0xFFFFFFFFu : ((1u << valid) - 1u);
      pop += __popc(ctr[q] & m);
    }
  }
  for (int off = 32; off >= 1; off >>= 1) pop += __shfl_xor(pop, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = pop;
  __syncthreads();
  if (threadIdx.x == 0) sums[(long)t * nconf + c] = 2 * (red[0] + red[1] + red[2] + red[3]) - vol;
}

// ------------------------------------------------------------------ memory network
// A pointwise proposal network over each site's own past field.  Every site carries two
// decayed memories M_j = 1024 m_j, m_j(t) = sum_{k >= 1} rho_j^(k-1) phi(t - k), in int16
// fixed point (M_j(0) = 0, M_j(t + 1) = floor((r_j M_j + 2^14) / 2^15) + 1024 phi(t), r_j =
// round(2^15 rho_j), rho = 0.75 and 0.94), packed in one word per site (M_0 low, M_1 high).
// With s_j = M_j / 1024 the logit is
//   a_k = max((W1[k][0] s_0 + W1[k][1] s_1) + b1[k], 0)                    (k < 8)
//   l   = ((c_0 s_0 + c_1 s_1) + b2 + sum_k W2[k] a_k) 2^-10               (k in order)
// every operation one FP32 rounding, no contraction; parameters [W1 (16), b1 (8), W2 (8),
// c (2), b2].  Site i of slice t is +1 where output i % 4 of the Philox block (i / 4, t, c,
// stream 4) is below T = proposal_threshold(l): q(+1) = T / 2^32 exactly and the site's
// log(p / q) is log(2^31 / f), f = T or 2^32 - T.  The linear part alone is the fields with
// memory with two decays; the hidden units let the tilt bend with the memory.
constexpr int MN_H = 8;
constexpr int MN_NP = 35;
constexpr int MN_R0 = 24576;  // round(2^15 0.75)
constexpr int MN_R1 = 30802;  // round(2^15 0.94)
constexpr int kMnMaxVol = 32768;

__device__ __forceinline__ float memnet_logit(const float* P, float s0, float s1) {
#pragma clang fp contract(off)
  float acc = P[32] * s0;
  acc = acc + P[33] * s1;
  acc = acc + P[34];
#pragma unroll
  for (int k = 0; k < MN_H; ++k) {
    float z = P[2 * k] * s0;
    z = z + P[2 * k + 1] * s1;
    z = z + P[16 + k];
    acc = acc + P[24 + k] * fmaxf(z, 0.f);
  }
  return acc * 0x1p-10f;
}

// the logit and the hidden units' inputs, in memnet_logit's order and rounding
__device__ __forceinline__ float memnet_logit_pre(const float* P, float s0, float s1,
                                                  float* pre) {
#pragma clang fp contract(off)
  float acc = P[32] * s0;
  acc = acc + P[33] * s1;
  acc = acc + P[34];
#pragma unroll
  for (int k = 0; k < MN_H; ++k) {
    float z = P[2 * k] * s0;
    z = z + P[2 * k + 1] * s1;
    z = z + P[16 + k];
    pre[k] = z;
    acc = acc + P[24 + k] * fmaxf(z, 0.f);
  }
  return acc * 0x1p-10f;
}

// the site's two memories after a draw
__device__ __forceinline__ uint32_t memnet_next(uint32_t st, bool up) {
  const int m0 = (int16_t)(st & 0xFFFFu), m1 = (int16_t)(st >> 16);
  const int d = up ? 1024 : -1024;
  const int n0 = ((MN_R0 * m0 + (1 << 14)) >> 15) + d;
  const int n1 = ((MN_R1 * m1 + (1 << 14)) >> 15) + d;
  return (uint32_t)(uint16_t)(int16_t)n0 | ((uint32_t)(uint16_t)(int16_t)n1 << 16);
}

__device__ __forceinline__ float mn_s0(uint32_t st) {
  return (float)(int16_t)(st & 0xFFFFu) * (1.f / 1024.f);
}
__device__ __forceinline__ float mn_s1(uint32_t st) {
  return (float)(int16_t)(st >> 16) * (1.f / 1024.f);
}

// sum over the workgroup (NTH threads) in a fixed order
template <int NTH>
__device__ __forceinline__ double mn_block_sum(double v, double* red) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = (red[0] + red[1]) + (red[2] + red[3]);
  if constexpr (NTH == 512) s += (red[4] + red[5]) + (red[6] + red[7]);
  __syncthreads();
  return s;
}

// The sampler: slice t of configurations config0 .. config0 + nconf - 1, one workgroup per
// configuration, each thread whole words; state[c] holds the memories, 32 words per field
// word (not read at t = 0); lr[c] receives the slice's log(p / q).
__global__ __launch_bounds__(256) void k_z2_memnet(uint64_t seed, uint32_t config0, uint32_t t,
                                                   int vol, const float* __restrict__ prm,
                                                   uint32_t* __restrict__ words,
                                                   uint32_t* __restrict__ state, double* lr,
                                                   const uint64_t* par) {
  __shared__ double red[4];
  if (par) {
    seed = par[0];
    config0 = (uint32_t)par[1];
  }
  float P[MN_NP];
#pragma unroll
  for (int i = 0; i < MN_NP; ++i) P[i] = prm[i];
  const int c = blockIdx.x;
  const int nw = (vol + 31) / 32;
  RatioAcc ra;
  for (int w = threadIdx.x; w < nw; w += 256) {
    uint4* sp = reinterpret_cast<uint4*>(state + ((long)c * nw + w) * 32);
    uint32_t ms[32];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint4 v = t == 0 ? make_uint4(0u, 0u, 0u, 0u) : sp[k];
      ms[4 * k] = v.x;
      ms[4 * k + 1] = v.y;
      ms[4 * k + 2] = v.z;
      ms[4 * k + 3] = v.w;
    }
    const int valid = vol - 32 * w;
    uint32_t bits = 0;
#pragma unroll
    for (int blk = 0; blk < 8; ++blk) {
      uint32_t ctr[4] = {(uint32_t)(8 * w + blk), t, config0 + (uint32_t)c, NRC_STREAM_MEMNET};
      philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = 4 * blk + q;
        const uint32_t th = proposal_threshold(memnet_logit(P, mn_s0(ms[i]), mn_s1(ms[i])));
        const bool up = ctr[q] < th;
        bits |= (uint32_t)up << i;
        if (i < valid) ra.add(up ? th : 0u - th);
        ms[i] = memnet_next(ms[i], up);
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
      sp[k] = make_uint4(ms[4 * k], ms[4 * k + 1], ms[4 * k + 2], ms[4 * k + 3]);
    words[(long)c * nw + w] = bits;
  }
  const double s = mn_block_sum<256>(ra.value(), red);
  if (threadIdx.x == 0) lr[c] = s;
}

// The proposal on the plain sampler's fields (stream 0, word w of slice t = output w % 4 of the
// block (w / 4, t, c, 0)), one workgroup of NTH threads per configuration walking its memories
// over t in registers.  GRAD false: out[t][c] = the cumulative log(p / q) of slices 0 .. t - 1.
// GRAD true: g[c][j] = 2^-10 sum_t G[t][c] sum_sites (up - sigmoid(l)) dl / dP_j, with
// l / 2^-10 differentiated (the gradient of sum log q with q(+1) = sigmoid(l), the smooth form
// of T / 2^32); each slice's per-thread FP32 sums go through LDS to thread j, which adds their
// FP64 total times G[t][c] to parameter j's; G = 0 skips a slice's products.
template <int WPT, bool GRAD, int NTH>
__global__ __launch_bounds__(NTH) void k_z2_memnet_walk(uint64_t seed, uint32_t config0, int T,
                                                        int vol, const float* __restrict__ prm,
                                                        const double* __restrict__ G,
                                                        double* __restrict__ out, int nconf) {
  __shared__ double red[NTH / 64];
  __shared__ float gs[GRAD ? MN_NP : 1][GRAD ? NTH : 1];
  float P[MN_NP];
#pragma unroll
  for (int i = 0; i < MN_NP; ++i) P[i] = prm[i];
  const int c = blockIdx.x;
  const int nw = (vol + 31) / 32;
  uint32_t ms[WPT][32];
#pragma unroll
  for (int k = 0; k < WPT; ++k)
#pragma unroll
    for (int j = 0; j < 32; ++j) ms[k][j] = 0u;
  double run = 0.0, gacc = 0.0;
  for (int t = 0; t < T; ++t) {
    if constexpr (!GRAD) {
      if (threadIdx.x == 0) out[(long)t * nconf + c] = run;
      if (t == T - 1) break;
    }
    const double gt = GRAD ? G[(long)t * nconf + c] : 0.0;
    RatioAcc ra;
    float g32[GRAD ? MN_NP : 1];
#pragma unroll
    for (int j = 0; j < (GRAD ? MN_NP : 1); ++j) g32[j] = 0.f;
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
      const int w = threadIdx.x + NTH * k;
      if (w < nw) {
        uint32_t ctr[4] = {(uint32_t)(w >> 2), (uint32_t)t, config0 + (uint32_t)c, 0u};
        philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
        const int q = w & 3;
        const uint32_t word = q == 0 ? ctr[0] : q == 1 ? ctr[1] : q == 2 ? ctr[2] : ctr[3];
        const int valid = vol - 32 * w;
#pragma unroll
        for (int i = 0; i < 32; ++i) {
          const bool up = (word >> i) & 1u;
          const float s0 = mn_s0(ms[k][i]), s1 = mn_s1(ms[k][i]);
          if constexpr (GRAD) {
            if (gt != 0.0 && i < valid) {
              float pre[MN_H];
              const float l = memnet_logit_pre(P, s0, s1, pre);
              const float dl = (up ? 1.f : 0.f) - 1.f / (1.f + __expf(-l));
#pragma unroll
              for (int h = 0; h < MN_H; ++h) {
                const float dz = pre[h] > 0.f ? dl * P[24 + h] : 0.f;
                g32[2 * h] += dz * s0;
                g32[2 * h + 1] += dz * s1;
                g32[16 + h] += dz;
                g32[24 + h] += dl * fmaxf(pre[h], 0.f);
              }
              g32[32] += dl * s0;
              g32[33] += dl * s1;
              g32[34] += dl;
            }
          } else {
            if (i < valid) {
              const uint32_t th = proposal_threshold(memnet_logit(P, s0, s1));
              ra.add(up ? th : 0u - th);
            }
          }
          ms[k][i] = memnet_next(ms[k][i], up);
        }
      }
    }
    if constexpr (GRAD) {
      if (gt != 0.0) {  // uniform: one configuration per workgroup
#pragma unroll
        for (int j = 0; j < MN_NP; ++j) gs[j][threadIdx.x] = g32[j];
        __syncthreads();
        if (threadIdx.x < MN_NP) {
          double s = 0.0;
          for (int i = 0; i < NTH; ++i) s += (double)gs[threadIdx.x][i];
          gacc += gt * s;
        }
        __syncthreads();
      }
    } else {
      run += mn_block_sum<NTH>(ra.value(), red);
    }
  }
  if constexpr (GRAD) {
    if (threadIdx.x < MN_NP) out[(long)c * MN_NP + threadIdx.x] = gacc * 0x1p-10;
  }
}

hipError_t launch_z2_memnet(uint64_t seed, uint32_t config0, int nconf, uint32_t t, int vol,
                            const float* prm, uint32_t* words, uint32_t* state, double* lr,
                            hipStream_t s, const uint64_t* par) {
  if (nconf < 1) return hipSuccess;
  hipLaunchKernelGGL(k_z2_memnet, dim3((unsigned)nconf), dim3(256), 0, s, seed, config0, t, vol,
                     prm, words, state, lr, par);
  return hipGetLastError();
}

hipError_t launch_z2_memnet_walk(uint64_t seed, uint32_t config0, int nconf, int T, int vol,
                                 const float* prm, const double* G, double* out, bool grad,
                                 hipStream_t s) {
  if (nconf < 1 || T < 1) return hipSuccess;
  if (vol < 1 || ((vol + 31) / 32) * 32 > kMnMaxVol) return hipErrorInvalidValue;
  const int nw = (vol + 31) / 32;
  const dim3 g((unsigned)nconf);
#define NRC_MN_WALK(W, NT)                                                                    \
  if (grad)                                                                                   \
    hipLaunchKernelGGL((k_z2_memnet_walk<W, true, NT>), g, dim3(NT), 0, s, seed, config0, T,   \
                       vol, prm, G, out, nconf);                                              \
  else                                                                                        \
    hipLaunchKernelGGL((k_z2_memnet_walk<W, false, NT>), g, dim3(NT), 0, s, seed, config0, T,  \
                       vol, prm, G, out, nconf)
  if (nw <= 256) {
    NRC_MN_WALK(1, 256);
  } else if (nw <= 512) {
    NRC_MN_WALK(2, 256);
  } else {
    NRC_MN_WALK(2, 512);
  }
#undef NRC_MN_WALK
  return hipGetLastError();
}

}  // namespace nrc

// ------------------------------------------------------------------ C ABI
using namespace nrc;


struct nrc_proposal {
  nrc_context* ctx = nullptr;
  Lat g{};
  long max_batch = 0;
  float* p = nullptr;  // master parameters
  float* grad = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  __bf16* wb = nullptr;
  __bf16* w2t = nullptr;
  __bf16* A1 = nullptr;  // [max_batch][V][16]
  __bf16* A2 = nullptr;
  __bf16* dZ1 = nullptr;  // training only, allocated on the first training call
  __bf16* dZ2 = nullptr;
  float* logits = nullptr;  // [max_batch][V]
  float* partial = nullptr;
  int partial_rows = 0;
  double* loss = nullptr;
  int step = 0;
};

#define PR_HIP(x)                              \
  do {                                         \
    if ((x) != hipSuccess) return NRC_ERR_HIP; \
  } while (0)

static int blocks_for(long n, int threads = 256, long cap = 8192) {
  long b = (n + threads - 1) / threads;
  if (b > cap) b = cap;
  return (int)(b < 1 ? 1 : b);
}

static void pr_free(nrc_proposal* p) {
  void* bufs[] = {p->p, p->grad, p->m, p->v, p->wb, p->w2t, p->A1, p->A2, p->dZ1, p->dZ2,
                  p->logits, p->partial, p->loss};
  for (void* b : bufs)
    if (b) hipFree(b);
  delete p;
}

extern "C" int nrc_proposal_create(nrc_context* ctx, int L, int max_batch, nrc_proposal** out) {
  if (!ctx || !out || max_batch < 1 || L < 4 || L > 64) return NRC_ERR_ARG;
  const long V = (long)L * L * L;
  if (V % 64 != 0) return NRC_ERR_UNSUPPORTED;
  hipSetDevice(ctx_device(ctx));
  nrc_proposal* p = new nrc_proposal();
  p->ctx = ctx;
  p->g = Lat{L, V};
  p->max_batch = max_batch;
  const size_t act = (size_t)max_batch * V * PC * sizeof(__bf16);
  bool ok = hipMalloc(&p->p, P_N * 4) == hipSuccess && hipMalloc(&p->grad, P_N * 4) == hipSuccess &&
            hipMalloc(&p->m, P_N * 4) == hipSuccess && hipMalloc(&p->v, P_N * 4) == hipSuccess &&
            hipMalloc(&p->wb, P_N * 2) == hipSuccess &&
            hipMalloc(&p->w2t, PC * NTAP * PC * 2) == hipSuccess &&
            hipMalloc(&p->A1, act) == hipSuccess && hipMalloc(&p->A2, act) == hipSuccess &&
            hipMalloc(&p->logits, (size_t)max_batch * V * 4) == hipSuccess &&
            hipMalloc(&p->loss, sizeof(double)) == hipSuccess;
  if (!ok) {
    pr_free(p);
    return NRC_ERR_NOMEM;
  }
  *out = p;
  return NRC_OK;
}

extern "C" int nrc_proposal_destroy(nrc_proposal* p) {
  if (!p) return NRC_ERR_ARG;
  hipStreamSynchronize(ctx_stream(p->ctx));
  pr_free(p);
  return NRC_OK;
}

extern "C" int nrc_proposal_param_count(void) { return P_N; }

extern "C" int nrc_proposal_set_params(nrc_proposal* p, const float* params) {
  if (!p || !params) return NRC_ERR_ARG;
  hipStream_t s = ctx_stream(p->ctx);
  PR_HIP(hipMemcpyAsync(p->p, params, P_N * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fill_zero_f, dim3(blocks_for(P_N)), dim3(256), 0, s, p->m, P_N);
  hipLaunchKernelGGL(k_fill_zero_f, dim3(blocks_for(P_N)), dim3(256), 0, s, p->v, P_N);
  hipLaunchKernelGGL(k_refresh, dim3(blocks_for(P_N)), dim3(256), 0, s, p->p, p->wb, p->w2t);
  PR_HIP(hipGetLastError());
  p->step = 0;
  PR_HIP(hipStreamSynchronize(s));
  return NRC_OK;
}

extern "C" int nrc_proposal_get_params(nrc_proposal* p, float* params) {
  if (!p || !params) return NRC_ERR_ARG;
  hipStream_t s = ctx_stream(p->ctx);
  PR_HIP(hipMemcpyAsync(params, p->p, P_N * 4, hipMemcpyDeviceToHost, s));
  PR_HIP(hipStreamSynchronize(s));
  return NRC_OK;
}

static int pr_forward_hidden(nrc_proposal* p, const uint32_t* prev, const float* u, long B) {
  hipStream_t s = ctx_stream(p->ctx);
  const long n = B * p->g.V;
  const int nw = (int)(p->g.V / 32);
  hipLaunchKernelGGL(k_conv1, dim3(blocks_for(n)), dim3(256), 0, s, prev, nw, u, p->wb, p->A1,
                     p->g, B);
  const long tiles = n / 16;
  hipLaunchKernelGGL((k_conv16<true>), dim3(blocks_for(tiles * 64, 256, 4096)), dim3(256), 0, s,
                     p->A1, p->wb + P_W2, p->wb + P_B2, (const __bf16*)nullptr, p->A2, p->g, B);
  PR_HIP(hipGetLastError());
  return NRC_OK;
}

// The fused inference kernel where the lattice fits its LDS ring (L <= 24); returns
// NRC_ERR_UNSUPPORTED for other L, which take the layer-by-layer kernels.
template <int L, int MODE>
static int pr_infer_L(nrc_proposal* p, const uint32_t* prev, const float* u, int B, float* lg,
                      uint64_t seed, uint32_t config0, uint32_t t, uint32_t stream,
                      uint32_t* words, double* logr) {
  static int occ[64] = {0};  // resident workgroups per CU, per device
  const int dev = ctx_device(p->ctx);
  int& o = occ[dev & 63];
  if (o == 0) {
    PR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_infer<L, MODE>, FUSED_NT, 0));
    if (o < 1) o = 1;
  }
  long grid = (long)o * ctx_n_cu(p->ctx);
  if (grid > B) grid = B;
  hipLaunchKernelGGL((k_infer<L, MODE>), dim3((unsigned)grid), dim3(FUSED_NT), 0,
                     ctx_stream(p->ctx), prev, u, p->wb, B, lg, seed, config0, t, stream, words,
                     logr);
  PR_HIP(hipGetLastError());
  return NRC_OK;
}

template <int MODE>
static int pr_infer(nrc_proposal* p, const uint32_t* prev, const float* u, int B, float* lg,
                    uint64_t seed, uint32_t config0, uint32_t t, uint32_t stream,
                    uint32_t* words, double* logr) {
  switch (p->g.L) {
    case 4: return pr_infer_L<4, MODE>(p, prev, u, B, lg, seed, config0, t, stream, words, logr);
    case 8: return pr_infer_L<8, MODE>(p, prev, u, B, lg, seed, config0, t, stream, words, logr);
    case 12:
      return pr_infer_L<12, MODE>(p, prev, u, B, lg, seed, config0, t, stream, words, logr);
    case 16:
      return pr_infer_L<16, MODE>(p, prev, u, B, lg, seed, config0, t, stream, words, logr);
    case 20:
      return pr_infer_L<20, MODE>(p, prev, u, B, lg, seed, config0, t, stream, words, logr);
    case 24:
      return pr_infer_L<24, MODE>(p, prev, u, B, lg, seed, config0, t, stream, words, logr);
    default: return NRC_ERR_UNSUPPORTED;
  }
}

// logits[B][V] (device) of the fields prev[B][nw] (device; null: no previous slice) and the
// normalised log-weights u[B] (device)
extern "C" int nrc_proposal_forward(nrc_proposal* p, const uint32_t* prev_dev, const float* u_dev,
                                    int B, float* logits_dev) {
  if (!p || !u_dev || !logits_dev || B < 1 || B > p->max_batch) return NRC_ERR_ARG;
  const int fr = pr_infer<0>(p, prev_dev, u_dev, B, logits_dev, 0ull, 0u, 0u, 0u, nullptr, nullptr);
  if (fr != NRC_ERR_UNSUPPORTED) return fr;
  if (pr_forward_hidden(p, prev_dev, u_dev, B) != NRC_OK) return NRC_ERR_HIP;
  hipStream_t s = ctx_stream(p->ctx);
  hipLaunchKernelGGL((k_conv3<0>), dim3(blocks_for((long)B * p->g.V)), dim3(256), 0, s, p->A2,
                     p->wb, p->g, (long)B, logits_dev, 0ull, 0u, 0u, 0u, (uint32_t*)nullptr,
                     (double*)nullptr, (const uint32_t*)nullptr, (const float*)nullptr, 0.f,
                     (double*)nullptr);
  PR_HIP(hipGetLastError());
  return NRC_OK;
}

// phi(t) for configurations config0 .. config0 + B - 1 into words_dev[B][nw], and
// logr_dev[c] += sum over sites of log(q_uniform / q(phi)) (FP64, device)
extern "C" int nrc_proposal_sample(nrc_proposal* p, const uint32_t* prev_dev, const float* u_dev,
                                   int B, uint64_t seed, uint32_t config0, uint32_t t,
                                   uint32_t stream, uint32_t* words_dev, double* logr_dev) {
  if (!p || !u_dev || !words_dev || !logr_dev || B < 1 || B > p->max_batch) return NRC_ERR_ARG;
  const int fr = pr_infer<1>(p, prev_dev, u_dev, B, nullptr, seed, config0, t, stream, words_dev,
                             logr_dev);
  if (fr != NRC_ERR_UNSUPPORTED) return fr;
  if (pr_forward_hidden(p, prev_dev, u_dev, B) != NRC_OK) return NRC_ERR_HIP;
  hipStream_t s = ctx_stream(p->ctx);
  hipLaunchKernelGGL((k_conv3<1>), dim3(blocks_for((long)B * p->g.V)), dim3(256), 0, s, p->A2,
                     p->wb, p->g, (long)B, (float*)nullptr, seed, config0, t, stream, words_dev,
                     logr_dev, (const uint32_t*)nullptr, (const float*)nullptr, 0.f,
                     (double*)nullptr);
  PR_HIP(hipGetLastError());
  return NRC_OK;
}

// Gradient of the weighted negative log-likelihood
//   loss = sum_b w[b] sum_s -log q(target_b(s) | prev_b, u_b) / (V sum_b w[b])
// into grad_out (host, P_N floats; may be null) and loss_out; with lr > 0 an optimizer step
// follows: Adam (optimizer 0; beta 0.9, 0.999, eps 1e-8) or SGD with momentum 0.9
// (optimizer 1).  All inputs are device arrays of B samples.
extern "C" int nrc_proposal_train_step(nrc_proposal* p, const uint32_t* prev_dev,
                                       const uint32_t* target_dev, const float* u_dev,
                                       const float* w_dev, int B, double wsum, float lr,
                                       int optimizer, float* grad_out, double* loss_out) {
  if (!p || !target_dev || !u_dev || !w_dev || B < 1 || B > p->max_batch || !(wsum > 0.0) ||
      (optimizer != 0 && optimizer != 1))
    return NRC_ERR_ARG;
  hipStream_t s = ctx_stream(p->ctx);
  const long n = (long)B * p->g.V;
  const int nw = (int)(p->g.V / 32);
  const size_t act = (size_t)p->max_batch * p->g.V * PC * sizeof(__bf16);
  if (!p->dZ1) {
    if (hipMalloc(&p->dZ1, act) != hipSuccess || hipMalloc(&p->dZ2, act) != hipSuccess)
      return NRC_ERR_NOMEM;
  }
  const int rows = 512;
  if (!p->partial) {
    if (hipMalloc(&p->partial, (size_t)rows * (NTAP * PC * PC + PC) * 4) != hipSuccess)
      return NRC_ERR_NOMEM;
    p->partial_rows = rows;
  }
  if (pr_forward_hidden(p, prev_dev, u_dev, B) != NRC_OK) return NRC_ERR_HIP;
  hipLaunchKernelGGL(k_fill_zero_d, dim3(1), dim3(64), 0, s, p->loss, 1);
  const float scale = (float)(1.0 / (wsum * (double)p->g.V));
  hipLaunchKernelGGL((k_conv3<2>), dim3(blocks_for(n)), dim3(256), 0, s, p->A2, p->wb, p->g,
                     (long)B, p->logits, 0ull, 0u, 0u, 0u, (uint32_t*)nullptr, (double*)nullptr,
                     target_dev, w_dev, scale, p->loss);
  hipLaunchKernelGGL(k_conv3_bwd, dim3(blocks_for(n)), dim3(256), 0, s, p->logits, p->A2, p->wb,
                     p->dZ2, p->g, (long)B);
  hipLaunchKernelGGL((k_conv16<false>), dim3(blocks_for(n / 16 * 64, 256, 4096)), dim3(256), 0,
                     s, p->dZ2, p->w2t, (const __bf16*)nullptr, p->A1, p->dZ1, p->g, (long)B);
  const long per_block = (n + rows - 1) / rows;
  // layer 3
  hipLaunchKernelGGL((k_wgrad<3>), dim3(rows), dim3(448), 0, s, p->logits, (const __bf16*)nullptr,
                     p->A2, (const uint32_t*)nullptr, nw, (const float*)nullptr, p->g, (long)B,
                     per_block, p->partial, NTAP * PC + 1);
  hipLaunchKernelGGL(k_sum_rows, dim3(blocks_for(NTAP * PC + 1)), dim3(256), 0, s, p->partial,
                     rows, NTAP * PC + 1, p->grad + P_W3);
  // layer 2
  hipLaunchKernelGGL((k_wgrad<2>), dim3(rows), dim3(256), 0, s, (const float*)nullptr, p->dZ2,
                     p->A1, (const uint32_t*)nullptr, nw, (const float*)nullptr, p->g, (long)B,
                     per_block, p->partial, NTAP * PC * PC + PC);
  hipLaunchKernelGGL(k_sum_rows, dim3(blocks_for(NTAP * PC * PC + PC)), dim3(256), 0, s,
                     p->partial, rows, NTAP * PC * PC + PC, p->grad + P_W2);
  // layer 1
  hipLaunchKernelGGL((k_wgrad<1>), dim3(rows), dim3(448), 0, s, (const float*)nullptr, p->dZ1,
                     (const __bf16*)nullptr, prev_dev, nw, u_dev, p->g, (long)B, per_block,
                     p->partial, PC * NTAP * 2 + PC);
  hipLaunchKernelGGL(k_sum_rows, dim3(blocks_for(PC * NTAP * 2 + PC)), dim3(256), 0, s,
                     p->partial, rows, PC * NTAP * 2 + PC, p->grad + P_W1);
  PR_HIP(hipGetLastError());
  if (lr > 0.f) {
    ++p->step;
    if (optimizer == 1)
      hipLaunchKernelGGL(k_sgdm, dim3(blocks_for(P_N)), dim3(256), 0, s, p->p, p->grad, p->m,
                         P_N, lr);
    else
      hipLaunchKernelGGL(k_adam, dim3(blocks_for(P_N)), dim3(256), 0, s, p->p, p->grad, p->m,
                         p->v, P_N, lr, 0.9f, 0.999f, 1e-8f, p->step);
    hipLaunchKernelGGL(k_refresh, dim3(blocks_for(P_N)), dim3(256), 0, s, p->p, p->wb, p->w2t);
    PR_HIP(hipGetLastError());
  }
  if (grad_out) PR_HIP(hipMemcpyAsync(grad_out, p->grad, P_N * 4, hipMemcpyDeviceToHost, s));
  double l = 0.0;
  PR_HIP(hipMemcpyAsync(&l, p->loss, sizeof(double), hipMemcpyDeviceToHost, s));
  PR_HIP(hipStreamSynchronize(s));
  if (loss_out) *loss_out = l * (double)scale;
  return NRC_OK;
}

extern "C" int nrc_norm_logw(nrc_context* ctx, const double* logabs_dev, int stride, int B,
                             double mu, double sd, float* u_dev) {
  if (!ctx || !logabs_dev || !u_dev || B < 1 || stride < 1 || !(sd > 0.0)) return NRC_ERR_ARG;
  hipLaunchKernelGGL(k_norm_logw, dim3(blocks_for(B)), dim3(256), 0, ctx_stream(ctx), logabs_dev,
                     stride, B, mu, sd, u_dev);
  PR_HIP(hipGetLastError());
  return NRC_OK;
}

extern "C" int nrc_z2_bits_list(nrc_context* ctx, uint64_t seed, const uint32_t* configs_dev,
                                const uint32_t* ts_dev, int n, uint32_t stream, int vol,
                                uint32_t* words_dev) {
  if (!ctx || !configs_dev || !ts_dev || !words_dev || n < 0 || vol < 1) return NRC_ERR_ARG;
  if (n == 0) return NRC_OK;
  const int nw = (vol + 31) / 32;
  hipLaunchKernelGGL(k_z2_bits_list, dim3(blocks_for((long)n * ((nw + 3) / 4))), dim3(256), 0,
                     ctx_stream(ctx), seed, configs_dev, ts_dev, n, stream, nw, words_dev);
  PR_HIP(hipGetLastError());
  return NRC_OK;
}

extern "C" int nrc_z2_sums(nrc_context* ctx, uint64_t seed, uint32_t config0, int nconf, int T,
                           uint32_t stream, int vol, int* sums_dev) {
  if (!ctx || !sums_dev || nconf < 0 || T < 0 || vol < 1) return NRC_ERR_ARG;
  if (nconf == 0 || T == 0) return NRC_OK;
  if (nconf > 2147483647 / T) return NRC_ERR_ARG;
  hipLaunchKernelGGL(k_z2_sums, dim3((unsigned)nconf, (unsigned)T), dim3(256), 0, ctx_stream(ctx),
                     seed, config0, nconf, stream, vol, sums_dev);
  PR_HIP(hipGetLastError());
  return NRC_OK;
}
