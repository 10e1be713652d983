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
// K-space split (k-space storage): pre = post = 1/xi_half, mid = (1 + A phi)/V, gather
// before the slice (the measurement at t precedes the evolution, production/slater_det/
// main.C:206-215).  X-space split (x-space storage): mid = kappa/V, gather at mid (full
// forward transform), post = 1 + A phi.
//
// Thread t owns the x-pencils (y, z) = t + q*NT, q < PPT, so every global load and store is
// a coalesced run of consecutive (y, z).  With PPT = NH each thread also owns exactly one row
// and one column of every LDS round.  L = 32 FP32: 512 threads, two pencils per thread, two
// rounds of 16 planes in 135 KB of LDS; L <= 24 FP32 and L <= 16 FP64: one round.
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
template <typename T>
__global__ void k_gather_sites(const C2<T>* psi, const int* site_of_sink, double2* M, long C,
                               int N, long V) {
  const long n = C * N * N;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long c = i / ((long)N * N);
    const int r = (int)(i - c * N * N), s = r / N, o = r - (r / N) * N;
    const C2<T> val = psi[(c * N + o) * V + site_of_sink[s]];
    M[i] = make_double2((double)val.x, (double)val.y);
  }
}

hipError_t launch_gather_sites(int prec, const void* psi, const int* site_of_sink, double2* M,
                               long C, int N, long V, hipStream_t s) {
  const long n = C * N * N;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  if (prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_gather_sites<double>, dim3((unsigned)blocks), dim3(256), 0, s,
                       (const C2<double>*)psi, site_of_sink, M, C, N, V);
  else
    hipLaunchKernelGGL(k_gather_sites<float>, dim3((unsigned)blocks), dim3(256), 0, s,
                       (const C2<float>*)psi, site_of_sink, M, C, N, V);
  return hipGetLastError();
}

template <typename T, int L, int PPT>
static hipError_t launch_s1(const SliceArgs& a, hipStream_t s) {
  const dim3 grid((unsigned)a.nvol), block(L * L / PPT);
  // k-space split: 1/xi_half, field, 1/xi_half, measure first; x-space split: kappa/V in
  // the middle with the measurement there, the field last
  if (a.gather_pre) return hipErrorInvalidValue;  // measured by k_gather_sites beforehand
  if (a.pre.kind == MUL_TABLE && a.mid.kind == MUL_FIELD && a.post.kind == MUL_TABLE &&
      !a.gather_mid)
    hipLaunchKernelGGL((k_slice1<T, L, PPT, MUL_TABLE, MUL_FIELD, MUL_TABLE, false, false>),
                       grid, block, 0, s, a);
  else if (a.pre.kind == MUL_TABLE && a.mid.kind == MUL_RFIELD && a.post.kind == MUL_TABLE &&
           !a.gather_mid)
    hipLaunchKernelGGL((k_slice1<T, L, PPT, MUL_TABLE, MUL_RFIELD, MUL_TABLE, false, false>),
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
