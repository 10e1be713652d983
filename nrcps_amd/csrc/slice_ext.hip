// slice_ext.hip -- the k-space split slice loop with FP32 state and an FP64 x-chain.
//
// The state of an orbital lives in FP32 (HBM, registers, LDS), but the part of every slice
// where FP32 rounding reaches the determinant is evaluated in FP64 registers.  Rounding
// errors matter in proportion to the size of the values they are relative to, seen from the
// sink momenta: an error in k-space at a sink momentum enters log|det| with weight one, one
// made after the x-transform (in (x, k_y, k_z) coordinates, where a sink's plane wave is
// spread over L sites) with weight 1/sqrt(L), and errors in the y and z transforms and the
// x-space field multiply with weights 1/L and below.  So per x-pencil the chain
//
//      x-DFT(+1) [finishes slice t-1] -> gather the sinks (measurement at t)
//      -> kinetic factor K -> x-DFT(-1) [starts slice t]
//
// runs on FP64 values converted from and rounded back to FP32 once, and between launches
// the state is kept in HBM in (x, k_y, k_z) coordinates with that x-DFT(+1) still pending.
// The y and z transforms and the field multiply stay in FP32 as in slice_fused.hip.  The
// FP64 FMA issues at the rate of the unpacked FP32 one on CDNA4, so the FP64 chain costs
// about what the compensated FP32 twiddles and factors it replaces cost.
//
// One workgroup per CU walks the volumes (vol = blockIdx.x, + gridDim.x, ...); a volume
// gets `nslices` consecutive slices while it stays on chip.  Launch 0 of a run starts from
// the k-space sources (start_k: no x-DFT(+1), factor K(1/2), unscaled gather); the last
// launch finishes with the x-DFT(+1) and stores the folded k-space state (store_k), the
// state every other entry point expects.
#include <type_traits>

#include "nrc_ops.h"

namespace nrc {

template <typename T, int L>
__device__ __forceinline__ void mul_field_lds_x(C2<T>* w, const uint32_t* bl, uint32_t off, T ph,
                                                T pl, T mh, T ml) {
#pragma unroll
  for (int y = 0; y < L; ++y) {
    const uint32_t site = off + (uint32_t)(y * L);
    const bool up = (bl[site >> 5] >> (site & 31u)) & 1u;
    w[y] = scale2<T>(w[y], up ? ph : mh, up ? pl : ml);
  }
}

// the z-DFT of one LDS row in place: FP64 registers rounded once (ZD), or FP32; Q: then
// times q[z]
template <int L, int S, bool ZD, bool Q>
__device__ __forceinline__ void z_row(C2<float>* row, const double* q) {
  if constexpr (ZD) {
    C2<double> w[L];
#pragma unroll
    for (int z = 0; z < L; ++z) w[z] = C2<double>{(double)row[z].x, (double)row[z].y};
    fft_reg<double, L, S>(w);
#pragma unroll
    for (int z = 0; z < L; ++z) {
      if constexpr (Q) w[z] = C2<double>{w[z].x * q[z], w[z].y * q[z]};
      row[z] = C2<float>{(float)w[z].x, (float)w[z].y};
    }
  } else {
    C2<float> w[L];
#pragma unroll
    for (int z = 0; z < L; ++z) w[z] = row[z];
    fft_reg<float, L, S>(w);
#pragma unroll
    for (int z = 0; z < L; ++z) {
      if constexpr (Q) {
        const double f = q[z];
        const float h = (float)f, l = (float)(f - (double)h);
        w[z] = scale2<float>(w[z], h, l);
      }
      row[z] = w[z];
    }
  }
}

// w[y] *= (a0 + a1 phi(site)) qy[y]: the factor formed in FP64 and applied as a compensated
// (hi, lo) pair
template <int L>
__device__ __forceinline__ void mul_field_lds_q(C2<float>* w, const uint32_t* bl, uint32_t off,
                                                double dp, double dm, const double* qy) {
#pragma unroll
  for (int y = 0; y < L; ++y) {
    const uint32_t site = off + (uint32_t)(y * L);
    const bool up = (bl[site >> 5] >> (site & 31u)) & 1u;
    const double f = (up ? dp : dm) * qy[y];
    const float h = (float)f, l = (float)(f - (double)h);
    w[y] = scale2<float>(w[y], h, l);
  }
}

// SNAP: after the x-DFT(+1) that finishes slice sl - 1 of the launch (sl >= 1), the folded
// k-space state is also written to snap + (sl - 1) snap_stride (FP32, the volume layout), for
// measurements made outside the launch (dense sinks, pair sinks)
template <int L, int PPT, int SMAX, bool ZD, bool TRAP, bool SNAP>
__global__ __launch_bounds__(L * L / PPT) void k_slicex(SliceXArgs a) {
  using C = C2<float>;
  using D = C2<double>;
  constexpr int P = L * L, NT = P / PPT, NH = PPT, H = L / NH, LP = L + 1;
  constexpr long V = (long)L * P;
  constexpr int NW = (int)(V / 32), WPT = (NW + NT - 1) / NT;
  static_assert(H * NH == L && NT * PPT == P && V % 32 == 0, "shape");
  __shared__ C sm[H * L * LP];
  __shared__ uint32_t bl[SMAX][NW];
  const int tid = threadIdx.x;
  const int rx = tid / L, rq = tid - (tid / L) * L;
  const int ns = a.nslices < SMAX ? a.nslices : SMAX;
  const double dp = a.a0 + a.a1, dm = a.a0 - a.a1;
  const float ph = (float)dp, mh = (float)dm;
  const float pl = (float)(dp - (double)ph), ml = (float)(dm - (double)mh);
  const rsrc_t rkx = make_rsrc(a.tab_x, (uint32_t)(V * sizeof(double)));
  const rsrc_t rkk = make_rsrc(a.tab_k, (uint32_t)(V * sizeof(double)));
  // separable trap q = qx(x) qy(y) qz(z): qx after the x-DFT(-1) (x in position space), qz
  // after the z-DFT(-1) rows, qy with the field; the index is the register index, the same in
  // every lane
  const double* qx = a.trap_sep;
  const double* qy = a.trap_sep + L;
  const double* qz = a.trap_sep + 2 * L;
  C* psi = reinterpret_cast<C*>(a.psi);
  for (long vol = blockIdx.x; vol < a.nvol; vol += gridDim.x) {
    const long cfg = vol / a.norb;
    const int orb = (int)(vol - cfg * a.norb);
    const rsrc_t rv = make_rsrc(psi + vol * V, (uint32_t)(V * sizeof(C)));
    // the configuration's field words of every slice of the launch, then the pencils
    uint32_t wb[SMAX][WPT];
#pragma unroll
    for (int sl = 0; sl < SMAX; ++sl) {
      const bool ok = sl < ns;
      const rsrc_t rb = make_rsrc(a.bits + (ok ? sl : 0) * a.bits_stride + cfg * a.nw,
                                  ok ? (uint32_t)NW * 4u : 0u);
#pragma unroll
      for (int k = 0; k < WPT; ++k)
        wb[sl][k] = bload<uint32_t>(rb, (uint32_t)(tid + k * NT) * 4u, 0u);
    }
    C v[PPT][L];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      const uint32_t yz = (uint32_t)(tid + q * NT);
#pragma unroll
      for (int x = 0; x < L; ++x)
        v[q][x] = bload<C>(rv, yz * (uint32_t)sizeof(C), (uint32_t)(x * P * sizeof(C)));
    }
    // the previous volume's last LDS reads are behind the barrier that ended its last round
#pragma unroll
    for (int sl = 0; sl < SMAX; ++sl)
#pragma unroll
      for (int k = 0; k < WPT; ++k)
        if (tid + k * NT < NW) bl[sl][tid + k * NT] = wb[sl][k];
    for (int sl = 0; sl < ns; ++sl) {
      const bool first = a.start_k && sl == 0;  // no x-DFT(+1): the state is in k-space
      const bool half = a.half0 && sl == 0;     // the sources: K(1/2), unscaled gather
      GatherSpec g = a.g;
      g.M = a.g.M + (long)sl * a.m_stride;
      g.sink_scale = half ? nullptr : a.g.sink_scale;
      // the x-chain in FP64, pencil by pencil; with two pencils per thread the idle one is
      // parked in LDS (free between rounds; thread-private slots [x][tid]) so that only one
      // FP64 pencil is live at a time
      auto chain = [&](D* w, uint32_t yz) {
        if (!first) fft_reg<double, L, +1>(w);
        if constexpr (SNAP) {
          if (sl > 0) {
            const rsrc_t rs = make_rsrc(reinterpret_cast<C*>(a.snap) + (sl - 1) * a.snap_stride +
                                            vol * V,
                                        (uint32_t)(V * sizeof(C)));
#pragma unroll
            for (int x = 0; x < L; ++x)
              bstore<C>(C{(float)w[x].x, (float)w[x].y}, rs, yz * (uint32_t)sizeof(C),
                        (uint32_t)(x * P * sizeof(C)));
          }
        }
        if (a.gather) gather_pencil<double, L>(w, g, (int)yz, cfg, orb, a.norb);
        const rsrc_t rk = half ? rkx : rkk;
#pragma unroll
        for (int x0 = 0; x0 < L; x0 += 8) {
          double f[8];
#pragma unroll
          for (int x = x0; x < x0 + 8 && x < L; ++x)
            f[x - x0] = bload<double>(rk, yz * 8u, (uint32_t)(x * P * 8));
#pragma unroll
          for (int x = x0; x < x0 + 8 && x < L; ++x)
            w[x] = D{w[x].x * f[x - x0], w[x].y * f[x - x0]};
        }
        fft_reg<double, L, -1>(w);
        if constexpr (TRAP) {
#pragma unroll
          for (int x = 0; x < L; ++x) w[x] = D{w[x].x * qx[x], w[x].y * qx[x]};
        }
      };
      if constexpr (PPT == 1) {
        D w[L];
#pragma unroll
        for (int x = 0; x < L; ++x) w[x] = D{(double)v[0][x].x, (double)v[0][x].y};
        chain(w, (uint32_t)tid);
#pragma unroll
        for (int x = 0; x < L; ++x) v[0][x] = C{(float)w[x].x, (float)w[x].y};
      } else {
        static_assert(PPT == 2 && L * NT <= H * L * LP, "parking slots");
#pragma unroll
        for (int x = 0; x < L; ++x) sm[x * NT + tid] = v[1][x];
        {
          D w[L];
#pragma unroll
          for (int x = 0; x < L; ++x) w[x] = D{(double)v[0][x].x, (double)v[0][x].y};
          chain(w, (uint32_t)tid);
#pragma unroll
          for (int x = 0; x < L; ++x) v[0][x] = C{(float)w[x].x, (float)w[x].y};
        }
        __builtin_amdgcn_sched_barrier(0);
        {
          D w[L];
#pragma unroll
          for (int x = 0; x < L; ++x) {  // swap: pencil 1 out of its slots, pencil 0 in
            const C t = sm[x * NT + tid];
            sm[x * NT + tid] = v[0][x];
            w[x] = D{(double)t.x, (double)t.y};
          }
          chain(w, (uint32_t)(tid + NT));
#pragma unroll
          for (int x = 0; x < L; ++x) v[1][x] = C{(float)w[x].x, (float)w[x].y};
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int x = 0; x < L; ++x) v[0][x] = sm[x * NT + tid];
        __syncthreads();  // every slot read before the first round writes its planes
      }
      // the y and z transforms with the field, in FP32, through LDS in NH rounds
      const uint32_t* b = bl[sl];
#pragma unroll
      for (int h = 0; h < NH; ++h) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
          const int yz = tid + q * NT, y = yz / L, z = yz - (yz / L) * L;
#pragma unroll
          for (int xl = 0; xl < H; ++xl) sm[(xl * L + y) * LP + z] = v[q][h * H + xl];
        }
        __syncthreads();
        z_row<L, -1, ZD, TRAP>(sm + (rx * L + rq) * LP, qz);
        __syncthreads();
        {  // columns along y with the field
          C w[L];
          C* col = sm + rx * L * LP + rq;
#pragma unroll
          for (int y = 0; y < L; ++y) w[y] = col[y * LP];
          fft_reg<float, L, -1>(w);
          const int x = h * H + rx;
          if constexpr (TRAP)
            mul_field_lds_q<L>(w, b, (uint32_t)(x * P + rq), dp, dm, qy);
          else
            mul_field_lds_x<float, L>(w, b, (uint32_t)(x * P + rq), ph, pl, mh, ml);
          fft_reg<float, L, +1>(w);
#pragma unroll
          for (int y = 0; y < L; ++y) col[y * LP] = w[y];
        }
        __syncthreads();
        z_row<L, +1, ZD, false>(sm + (rx * L + rq) * LP, qz);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
          const int yz = tid + q * NT, y = yz / L, z = yz - (yz / L) * L;
#pragma unroll
          for (int xl = 0; xl < H; ++xl) v[q][h * H + xl] = sm[(xl * L + y) * LP + z];
        }
        __syncthreads();  // the planes are free for the next round, slice or volume
      }
    }
    // store: the pending x-DFT(+1) left for the next launch, or done here (last launch)
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      const uint32_t yz = (uint32_t)(tid + q * NT);
      if (a.store_k) {
        D w[L];
#pragma unroll
        for (int x = 0; x < L; ++x) w[x] = D{(double)v[q][x].x, (double)v[q][x].y};
        fft_reg<double, L, +1>(w);
#pragma unroll
        for (int x = 0; x < L; ++x) v[q][x] = C{(float)w[x].x, (float)w[x].y};
      }
#pragma unroll
      for (int x = 0; x < L; ++x)
        bstore<C>(v[q][x], rv, yz * (uint32_t)sizeof(C), (uint32_t)(x * P * sizeof(C)));
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

template <int L, int PPT, bool ZD, bool TRAP, bool SNAP>
static hipError_t launch_sx_s(const SliceXArgs& a, hipStream_t s) {
  constexpr int SMAX = SLICE_FUSE_MAX;
  static int per_cu = 0;
  if (per_cu == 0) {
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, k_slicex<L, PPT, SMAX, ZD, TRAP, SNAP>, L * L / PPT, 0);
    per_cu = (e != hipSuccess || nb < 1) ? 1 : nb;
  }
  long grid = (long)a.grid * per_cu;
  if (grid > a.nvol) grid = a.nvol;
  hipLaunchKernelGGL((k_slicex<L, PPT, SMAX, ZD, TRAP, SNAP>), dim3((unsigned)grid),
                     dim3(L * L / PPT), 0, s, a);
  return hipGetLastError();
}

template <int L, int PPT, bool ZD, bool TRAP>
static hipError_t launch_sx_t(const SliceXArgs& a, hipStream_t s) {
  return a.snap && a.nslices > 1 ? launch_sx_s<L, PPT, ZD, TRAP, true>(a, s)
                                 : launch_sx_s<L, PPT, ZD, TRAP, false>(a, s);
}

// ZD: the shape may run its z rows in FP64 (a.zd chooses); false: FP32 z rows only
template <int L, int PPT, bool ZD = true>
static hipError_t launch_sx(const SliceXArgs& a, hipStream_t s) {
  if (a.nslices < 1 || a.nslices > SLICE_FUSE_MAX) return hipErrorInvalidValue;
  if (ZD && a.zd)
    return a.trap_sep ? launch_sx_t<L, PPT, true, true>(a, s)
                      : launch_sx_t<L, PPT, true, false>(a, s);
  return a.trap_sep ? launch_sx_t<L, PPT, false, true>(a, s)
                    : launch_sx_t<L, PPT, false, false>(a, s);
}

bool slicex_supported(int L) { return L == 8 || L == 12 || L == 16 || L == 24 || L == 32; }

// tiles (x-pencils per thread) a shape offers: the first is the default
int slicex_tiles(int L, int* tiles) {
  if (L == 32) {  // one pencil per thread would need twice the registers
    tiles[0] = 2;
    return 1;
  }
  if (!slicex_supported(L)) return 0;
  tiles[0] = 1;
  tiles[1] = 2;
  return 2;
}

hipError_t launch_slicex(const SliceXArgs& a, hipStream_t s) {
  if (a.nvol <= 0) return hipSuccess;
  if (a.grid < 1 || !a.tab_x || !a.tab_k || !a.bits) return hipErrorInvalidValue;
  const bool two = a.tile == 2;
  if (a.tile != 0 && a.tile != 1 && a.tile != 2) return hipErrorInvalidValue;
  switch (a.L) {
    case 8: return two ? launch_sx<8, 2>(a, s) : launch_sx<8, 1>(a, s);
    case 12: return two ? launch_sx<12, 2>(a, s) : launch_sx<12, 1>(a, s);
    case 16: return two ? launch_sx<16, 2>(a, s) : launch_sx<16, 1>(a, s);
    // FP64 rows spill at this shape
    case 24: return two ? launch_sx<24, 2, false>(a, s) : launch_sx<24, 1, false>(a, s);
    case 32: return a.tile == 1 ? hipErrorInvalidValue : launch_sx<32, 2>(a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace nrc
