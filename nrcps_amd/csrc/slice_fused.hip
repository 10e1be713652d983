// slice_fused.hip -- one time slice of every orbital in a single pass over HBM.
//
// One workgroup evolves one orbital volume (L^3 complex) through a whole split-operator
// slice, reading and writing each element exactly once:
//
//   load x-pencils -> [gather: measure] -> pre multiplier -> x-DFT (sign -1, registers)
//   for each of NH rounds of H = L/NH x-planes:
//       planes -> LDS; z-DFT (-1) per row; y-DFT (-1) per column, [gather], mid multiplier,
//       y-DFT (+1); z-DFT (+1) per row; LDS -> registers
//   x-DFT (+1, registers) -> post multiplier -> store
//
// K-space split (k-space storage): a slice is K(1/2) F+[(1 + A phi)/V F-[K(1/2) psi]]; the
// trailing K(1/2) of slice t and the leading one of slice t+1 are applied as one factor
// K = 1/xi_half^2 (pre; the first slice uses K(1/2)), no post multiplier, and the
// measurement, which precedes the evolution (production/slater_det/main.C:206-215), scales
// each gathered sink value by K(1/2)(k_s).  X-space split (x-space storage): mid = kappa/V,
// gather at mid (full forward transform), post = 1 + A phi.
//
// Thread t owns the x-pencils (y, z) = t + q*NT, q < PPT, so every global load and store is
// a coalesced run of consecutive (y, z).  With PPT = NH each thread also owns exactly one row
// and one column of every LDS round.  L = 32 FP32: 512 threads, two pencils per thread, two
// rounds of 16 planes in 135 KB of LDS; L <= 24 FP32 and L <= 16 FP64: one round.
#include <type_traits>

#include "nrc_ops.h"

namespace nrc {

template <typename T, int L, int PPT, int KPRE, int KMID, int KPOST, bool GPRE, bool GMID>
__global__ __launch_bounds__(L * L / PPT) void k_slice1(SliceArgs a) {
  using C = C2<T>;
  constexpr int P = L * L, NT = P / PPT, NH = PPT, H = L / NH, LP = L + 1;
  constexpr long V = (long)L * P;
  static_assert(H * NH == L && NT * PPT == P, "shape");
  __shared__ C sm[H * L * LP];
  const long vol = blockIdx.x;
  const long cfg = vol / a.norb;
  const int orb = (int)(vol - cfg * a.norb);
  const int tid = threadIdx.x;
  // this volume through a buffer resource: one lane offset per pencil, x * P in the constant
  const rsrc_t rv = make_rsrc(reinterpret_cast<C*>(a.psi) + vol * V, (uint32_t)(V * sizeof(C)));
  C v[PPT][L];
  // pencil by pencil (scheduling barriers keep the next pencil's loads and table reads from
  // being hoisted over this pencil's transform, which would not fit the register file)
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const int yz = tid + q * NT;
#pragma unroll
    for (int x = 0; x < L; ++x)
      v[q][x] = bload<C>(rv, (uint32_t)yz * (uint32_t)sizeof(C), (uint32_t)(x * P * sizeof(C)));
    if constexpr (GPRE) gather_pencil<T, L>(v[q], a.g, yz, cfg, orb, a.norb);
    apply_mul_k<T, L, P, KPRE>(v[q], a.pre, cfg, (uint32_t)yz, V);
    fft_reg<T, L, -1>(v[q]);
    __builtin_amdgcn_sched_barrier(0);
  }
  const int rx = tid / L, rq = tid - (tid / L) * L;  // this thread's (plane, row|column)
#pragma unroll
  for (int h = 0; h < NH; ++h) {
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      const int yz = tid + q * NT, y = yz / L, z = yz - (yz / L) * L;
#pragma unroll
      for (int xl = 0; xl < H; ++xl) sm[(xl * L + y) * LP + z] = v[q][h * H + xl];
    }
    __syncthreads();
    {  // rows along z
      C w[L];
      C* row = sm + (rx * L + rq) * LP;
#pragma unroll
      for (int z = 0; z < L; ++z) w[z] = row[z];
      fft_reg<T, L, -1>(w);
#pragma unroll
      for (int z = 0; z < L; ++z) row[z] = w[z];
    }
    __syncthreads();
    {  // columns along y, with the mid-slice gather and multiplier
      C w[L];
      C* col = sm + rx * L * LP + rq;
#pragma unroll
      for (int y = 0; y < L; ++y) w[y] = col[y * LP];
      fft_reg<T, L, -1>(w);
      const int x = h * H + rx;
      if constexpr (GMID) gather_pencil<T, L>(w, a.g, x * L + rq, cfg, orb, a.norb);
      apply_mul_k<T, L, L, KMID>(w, a.mid, cfg, (uint32_t)(x * P + rq), V);
      fft_reg<T, L, +1>(w);
#pragma unroll
      for (int y = 0; y < L; ++y) col[y * LP] = w[y];
    }
    __syncthreads();
    {  // rows back along z
      C w[L];
      C* row = sm + (rx * L + rq) * LP;
#pragma unroll
      for (int z = 0; z < L; ++z) w[z] = row[z];
      fft_reg<T, L, +1>(w);
#pragma unroll
      for (int z = 0; z < L; ++z) row[z] = w[z];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      const int yz = tid + q * NT, y = yz / L, z = yz - (yz / L) * L;
#pragma unroll
      for (int xl = 0; xl < H; ++xl) v[q][h * H + xl] = sm[(xl * L + y) * LP + z];
    }
    if (h + 1 < NH) __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const int yz = tid + q * NT;
    fft_reg<T, L, +1>(v[q]);
    apply_mul_k<T, L, P, KPOST>(v[q], a.post, cfg, (uint32_t)yz, V);
#pragma unroll
    for (int x = 0; x < L; ++x)
      bstore<C>(v[q][x], rv, (uint32_t)yz * (uint32_t)sizeof(C), (uint32_t)(x * P * sizeof(C)));
    __builtin_amdgcn_sched_barrier(0);
  }
}

// M[c][s][n] = psi[c][n][site_of_sink[s]]: the measurement of plane-wave sinks on orbitals
// stored in k-space (OneBody::Project of a unit vector, src/one_body/noarch/
// one_body_project.C:14-27), read straight from memory before the slice overwrites them.
// With conjugate partners (cp.vol != null) a configuration holds cp.nvo volumes and column n
// is volume cp.vol[n], read at the mirrored sink site cp.msite[s] and conjugated when
// cp.conj[n] != 0.
template <typename T>
__global__ void k_gather_sites(const C2<T>* psi, const int* site_of_sink, const double* scale,
                               double2* M, long C, int N, long V, ConjCols cp) {
  const long n = C * N * N;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long c = i / ((long)N * N);
    const int r = (int)(i - c * N * N), s = r / N, o = r - (r / N) * N;
    const bool cj = cp.vol && cp.conj[o];
    const long vol = cp.vol ? c * cp.nvo + cp.vol[o] : c * N + o;
    const C2<T> val = psi[vol * V + (cj ? cp.msite[s] : site_of_sink[s])];
    const double f = scale ? scale[s] : 1.0;
    M[i] = make_double2((double)val.x * f, (cj ? -(double)val.y : (double)val.y) * f);
  }
}

// M[c][s][n] = sink_scale[s] psi[c][n][site_of_sink[s]] (sink_scale null: 1)
hipError_t launch_gather_sites(int prec, const void* psi, const int* site_of_sink,
                               const double* sink_scale, double2* M, long C, int N, long V,
                               hipStream_t s, const ConjCols& cp) {
  const long n = C * N * N;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  if (prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_gather_sites<double>, dim3((unsigned)blocks), dim3(256), 0, s,
                       (const C2<double>*)psi, site_of_sink, sink_scale, M, C, N, V, cp);
  else
    hipLaunchKernelGGL(k_gather_sites<float>, dim3((unsigned)blocks), dim3(256), 0, s,
                       (const C2<float>*)psi, site_of_sink, sink_scale, M, C, N, V, cp);
  return hipGetLastError();
}

// w[y] *= (a0 + a1 phi(site)), site = off + y L, phi from the configuration's field words in
// LDS; (ph, pl) and (mh, ml) are the compensated factors for phi = +1 and -1
template <typename T, int L>
__device__ __forceinline__ void mul_field_lds(C2<T>* w, const uint32_t* bl, uint32_t off, T ph,
                                              T pl, T mh, T ml) {
#pragma unroll
  for (int y = 0; y < L; ++y) {
    const uint32_t site = off + (uint32_t)(y * L);
    const bool up = (bl[site >> 5] >> (site & 31u)) & 1u;
    w[y] = scale2<T>(w[y], up ? ph : mh, up ? pl : ml);
  }
}

// Persistent, software-pipelined k_slice1 for the k-space split (kinetic factor, field, no
// trailing factor): one workgroup per CU walks the volumes blockIdx.x, + gridDim.x, ...  A
// volume is loaded, transformed and stored as in k_slice1, but
//  * the loads of the next volume's first pencil are issued before this volume's last yz
//    rounds (PF = 1), before its final x-transforms (PF = 0), half at each of those points
//    (PF = 3, L = 32: fewer registers live through the rounds) or once its first pencil is
//    stored (PF = 2), so they land while the CU computes; loads complete in order, so everything this
//    volume still waits for is issued first: the configurations' field words (staged in LDS)
//    go out before the other pencils, and the prefetched pencil is multiplied by its first
//    kinetic factor at the end of the previous volume;
//  * a.nslices (<= SMAX) consecutive slices run on the volume while it stays on chip: between
//    two of them the plane-wave sinks are gathered from registers into the next slice's
//    overlap matrix (a.g, M advanced by a.m_stride per slice), and slice s > 0 starts with
//    the factor a.pre_next.  HBM traffic per slice drops by that factor.
template <typename T, int L, int PPT, int PF, int SMAX, bool TRAP>
__global__ __launch_bounds__(L * L / PPT) void k_slice1p(SliceArgs a) {
  using C = C2<T>;
  using E = TabEntry<T>;
  constexpr int P = L * L, NT = P / PPT, NH = PPT, H = L / NH, LP = L + 1;
  constexpr long V = (long)L * P;
  constexpr int NW = (int)(V / 32), WPT = (NW + NT - 1) / NT;
  static_assert(H * NH == L && NT * PPT == P && V % 32 == 0, "shape");
  __shared__ C sm[H * L * LP];
  __shared__ uint32_t bl[SMAX][NW];
  const int tid = threadIdx.x;
  const int rx = tid / L, rq = tid - (tid / L) * L;
  const int ns = a.nslices < SMAX ? a.nslices : SMAX;
  const double dp = a.mid.a0 + a.mid.a1, dm = a.mid.a0 - a.mid.a1;
  const T ph = T(dp), mh = T(dm);
  const T pl = T(dp - (double)ph), ml = T(dm - (double)mh);
  const rsrc_t rt = make_rsrc(a.pre.table, (uint32_t)(V * sizeof(E)));
  // trap: the factor tables (a0 +- a1) q, phi = +1 first (unused without one)
  const rsrc_t rq_trap = make_rsrc(a.mid.table, TRAP ? (uint32_t)(2 * V * sizeof(E)) : 0u);
  MulSpec knext = a.pre;
  knext.table = a.pre_next;
  C* psi = reinterpret_cast<C*>(a.psi);
  long vol = blockIdx.x;
  if (vol >= a.nvol) return;
  C nx[L];  // the first pencil of the volume about to be processed
  // Loads of the first pencil of volume w; past the last volume the resource has no records
  // and the loads return zeros without touching memory, so the load is unconditional (a
  // conditional load would keep the previous values alive across the whole volume).
  auto prefetch = [&](long w, auto x0_tag, auto x1_tag) {  // elements [x0, x1) of the pencil
    constexpr int X0 = decltype(x0_tag)::value, X1 = decltype(x1_tag)::value;
    const bool ok = w < a.nvol;
    const rsrc_t rn = make_rsrc(psi + (ok ? w : 0) * V, ok ? (uint32_t)(V * sizeof(C)) : 0u);
#pragma unroll
    for (int x = X0; x < X1; ++x)
      nx[x] = bload<C>(rn, (uint32_t)tid * (uint32_t)sizeof(C), (uint32_t)(x * P * sizeof(C)));
  };
  using X0 = std::integral_constant<int, 0>;
  using XH = std::integral_constant<int, L / 2>;
  using XL = std::integral_constant<int, L>;
  // the first slice factor of that pencil (the table is the same for every volume); issued
  // once the pencil's loads are out, so that waiting for it waits for nothing later
  auto factor = [&]() {
    E tab0[L];
#pragma unroll
    for (int x = 0; x < L; ++x)
      tab0[x] = bload<E>(rt, (uint32_t)tid * (uint32_t)sizeof(E), (uint32_t)(x * P * sizeof(E)));
#pragma unroll
    for (int x = 0; x < L; ++x) nx[x] = scale_tab<T>(nx[x], tab0[x]);
  };
  prefetch(vol, X0{}, XL{});
  factor();
  for (; vol < a.nvol; vol += gridDim.x) {
    const long cfg = vol / a.norb;
    const int orb = (int)(vol - cfg * a.norb);
    const rsrc_t rv = make_rsrc(psi + vol * V, (uint32_t)(V * sizeof(C)));
    const long nv = vol + gridDim.x;
    C v[PPT][L];
    // field words of every slice first, then the other pencils
    uint32_t wb[SMAX][WPT];
#pragma unroll
    for (int sl = 0; sl < SMAX; ++sl) {
      const bool ok = sl < ns;
      const rsrc_t rb = make_rsrc(a.mid.bits + (ok ? sl : 0) * a.bits_stride + cfg * a.mid.nw,
                                  ok ? (uint32_t)NW * 4u : 0u);
#pragma unroll
      for (int k = 0; k < WPT; ++k)
        wb[sl][k] = bload<uint32_t>(rb, (uint32_t)(tid + k * NT) * 4u, 0u);
    }
#pragma unroll
    for (int q = 1; q < PPT; ++q) {
      const uint32_t yz = (uint32_t)(tid + q * NT);
#pragma unroll
      for (int x = 0; x < L; ++x)
        v[q][x] = bload<C>(rv, yz * (uint32_t)sizeof(C), (uint32_t)(x * P * sizeof(C)));
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int x = 0; x < L; ++x) v[0][x] = nx[x];
#pragma unroll
    for (int q = 1; q < PPT; ++q) {
      apply_mul_k<T, L, P, MUL_TABLE>(v[q], a.pre, cfg, (uint32_t)(tid + q * NT), V);
      __builtin_amdgcn_sched_barrier(0);
    }
    // stage the words; the barrier also orders the previous volume's last LDS reads before
    // this volume's first plane writes
#pragma unroll
    for (int sl = 0; sl < SMAX; ++sl)
#pragma unroll
      for (int k = 0; k < WPT; ++k)
        if (tid + k * NT < NW) bl[sl][tid + k * NT] = wb[sl][k];
    __syncthreads();
    // one slice on the volume in registers and LDS; LAST: the volume's final slice (the
    // prefetch point of PF = 1), otherwise followed by the gather of the next measurement
    auto slice = [&](int sl, auto last_tag) {
      constexpr bool LAST = decltype(last_tag)::value;
      if (sl > 0) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
          apply_mul_k<T, L, P, MUL_TABLE>(v[q], knext, cfg, (uint32_t)(tid + q * NT), V);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        fft_reg<T, L, -1>(v[q]);
        __builtin_amdgcn_sched_barrier(0);
      }
      if constexpr (LAST && (PF == 1 || PF == 3)) {  // PF 3: the first half only
        if constexpr (PF == 1) prefetch(nv, X0{}, XL{});
        else prefetch(nv, X0{}, XH{});
        __builtin_amdgcn_sched_barrier(0);
      }
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
        {  // rows along z
          C w[L];
          C* row = sm + (rx * L + rq) * LP;
#pragma unroll
          for (int z = 0; z < L; ++z) w[z] = row[z];
          fft_reg<T, L, -1>(w);
#pragma unroll
          for (int z = 0; z < L; ++z) row[z] = w[z];
        }
        __syncthreads();
        {  // columns along y with the field
          C w[L];
          C* col = sm + rx * L * LP + rq;
#pragma unroll
          for (int y = 0; y < L; ++y) w[y] = col[y * LP];
          fft_reg<T, L, -1>(w);
          const int x = h * H + rx;
          if constexpr (TRAP)
            mul_field_trap<T, L>(w, b, (uint32_t)(x * P + rq), rq_trap, (uint32_t)V);
          else
            mul_field_lds<T, L>(w, b, (uint32_t)(x * P + rq), ph, pl, mh, ml);
          fft_reg<T, L, +1>(w);
#pragma unroll
          for (int y = 0; y < L; ++y) col[y * LP] = w[y];
        }
        __syncthreads();
        {  // rows back along z
          C w[L];
          C* row = sm + (rx * L + rq) * LP;
#pragma unroll
          for (int z = 0; z < L; ++z) w[z] = row[z];
          fft_reg<T, L, +1>(w);
#pragma unroll
          for (int z = 0; z < L; ++z) row[z] = w[z];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
          const int yz = tid + q * NT, y = yz / L, z = yz - (yz / L) * L;
#pragma unroll
          for (int xl = 0; xl < H; ++xl) v[q][h * H + xl] = sm[(xl * L + y) * LP + z];
        }
        __syncthreads();  // the planes are free for the next round, slice or volume
      }
      if constexpr (LAST && (PF == 0 || PF == 3)) {  // PF 3: the second half
        if constexpr (PF == 0) prefetch(nv, X0{}, XL{});
        else prefetch(nv, XH{}, XL{});
        __builtin_amdgcn_sched_barrier(0);
      }
      if constexpr (!LAST) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
          fft_reg<T, L, +1>(v[q]);
          GatherSpec g = a.g;
          g.M = a.g.M + (long)sl * a.m_stride;
          gather_pencil<T, L>(v[q], g, tid + q * NT, cfg, orb, a.norb);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    };
    for (int sl = 0; sl + 1 < ns; ++sl) slice(sl, std::integral_constant<bool, false>{});
    slice(ns - 1, std::integral_constant<bool, true>{});
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      const uint32_t yz = (uint32_t)(tid + q * NT);
      fft_reg<T, L, +1>(v[q]);
#pragma unroll
      for (int x = 0; x < L; ++x)
        bstore<C>(v[q][x], rv, yz * (uint32_t)sizeof(C), (uint32_t)(x * P * sizeof(C)));
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (PF == 2) {  // after the first pencil is stored, into its registers
        if (q == 0) prefetch(nv, X0{}, XL{});
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    factor();
    __builtin_amdgcn_sched_barrier(0);
  }
}

// a.grid = CUs; the grid is as many workgroups as are resident at once
template <typename T, int L, int PPT, int PF>
static hipError_t launch_s1p(const SliceArgs& a, hipStream_t s) {
  constexpr int SMAX = SLICE_FUSE_MAX;
  if (a.nslices < 1 || a.nslices > SMAX) return hipErrorInvalidValue;
  static int per_cu[2] = {0, 0};
  const bool trap = a.mid.table != nullptr;
  if (per_cu[trap] == 0) {
    int nb = 0;
    const hipError_t e =
        trap ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                   &nb, k_slice1p<T, L, PPT, PF, SMAX, true>, L * L / PPT, 0)
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                   &nb, k_slice1p<T, L, PPT, PF, SMAX, false>, L * L / PPT, 0);
    per_cu[trap] = (e != hipSuccess || nb < 1) ? 1 : nb;
  }
  long grid = (long)a.grid * per_cu[trap];
  if (grid > a.nvol) grid = a.nvol;
  if (trap)
    hipLaunchKernelGGL((k_slice1p<T, L, PPT, PF, SMAX, true>), dim3((unsigned)grid),
                       dim3(L * L / PPT), 0, s, a);
  else
    hipLaunchKernelGGL((k_slice1p<T, L, PPT, PF, SMAX, false>), dim3((unsigned)grid),
                       dim3(L * L / PPT), 0, s, a);
  return hipGetLastError();
}

template <typename T, int L, int PPT>
static hipError_t launch_s1(const SliceArgs& a, hipStream_t s) {
  const dim3 grid((unsigned)a.nvol), block(L * L / PPT);
  // the k-space split with the plain field: the persistent pipelined kernel when the caller
  // gives a workgroup count (prefetch point: before the yz rounds, split around them at
  // L = 32, or after the first pencil's store where the registers are short)
  if (a.grid > 0 && a.pre.kind == MUL_TABLE && a.mid.kind == MUL_FIELD &&
      a.post.kind == MUL_NONE && !a.gather_mid &&
      (a.mid.table == nullptr || slice1p_trap_ok(sizeof(T) == 8 ? NRC_PREC_F64 : NRC_PREC_F32, L)))
    return launch_s1p<T, L, PPT, (sizeof(T) == 4 && L == 24) ? 2 : (sizeof(T) == 4 && L == 32) ? 3 : 1>(a, s);
  if (a.nslices != 1) return hipErrorInvalidValue;  // fused slices: persistent kernel only
  // k-space split: kinetic factor, field (or smeared field), measured beforehand; x-space
  // split: kappa/V in the middle with the measurement there, the field last
  if (a.gather_pre) return hipErrorInvalidValue;  // measured by k_gather_sites beforehand
  if (a.pre.kind == MUL_TABLE && a.mid.kind == MUL_FIELD && a.post.kind == MUL_NONE &&
      !a.gather_mid)
    hipLaunchKernelGGL((k_slice1<T, L, PPT, MUL_TABLE, MUL_FIELD, MUL_NONE, false, false>),
                       grid, block, 0, s, a);
  else if (a.pre.kind == MUL_TABLE && a.mid.kind == MUL_RFIELD && a.post.kind == MUL_NONE &&
           !a.gather_mid)
    hipLaunchKernelGGL((k_slice1<T, L, PPT, MUL_TABLE, MUL_RFIELD, MUL_NONE, false, false>),
                       grid, block, 0, s, a);
  else if (a.pre.kind == MUL_NONE && a.mid.kind == MUL_TABLE && a.post.kind == MUL_FIELD &&
           !a.gather_pre && a.gather_mid)
    hipLaunchKernelGGL((k_slice1<T, L, PPT, MUL_NONE, MUL_TABLE, MUL_FIELD, false, true>), grid,
                       block, 0, s, a);
  else if (a.pre.kind == MUL_NONE && a.mid.kind == MUL_TABLE && a.post.kind == MUL_FIELD &&
           !a.gather_pre && !a.gather_mid)  // x-space split measured by dense overlaps
    hipLaunchKernelGGL((k_slice1<T, L, PPT, MUL_NONE, MUL_TABLE, MUL_FIELD, false, false>),
                       grid, block, 0, s, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// the persistent kernel with a trap's factor tables: not for FP32 L = 32, whose registers
// are full without them (that shape runs trapped slices one per launch in k_slice1)
bool slice1p_trap_ok(int prec, int L) { return !(prec == NRC_PREC_F32 && L == 32); }

bool slice1_supported(int prec, int L) {
  if (prec == NRC_PREC_F64) return L == 8 || L == 12 || L == 16;
  return L == 8 || L == 12 || L == 16 || L == 24 || L == 32;
}

hipError_t launch_slice1(int prec, const SliceArgs& a, hipStream_t s) {
  if (a.nvol <= 0) return hipSuccess;
  if (prec == NRC_PREC_F64) {
    switch (a.L) {
      case 8: return launch_s1<double, 8, 1>(a, s);
      case 12: return launch_s1<double, 12, 1>(a, s);
      case 16: return launch_s1<double, 16, 1>(a, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (a.L) {
    case 8: return launch_s1<float, 8, 1>(a, s);
    case 12: return launch_s1<float, 12, 1>(a, s);
    case 16: return launch_s1<float, 16, 1>(a, s);
    case 24: return launch_s1<float, 24, 1>(a, s);
    case 32: return launch_s1<float, 32, 2>(a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace nrc
