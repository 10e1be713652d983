// fft_passes.hip -- the two fused passes that make up one split-operator time slice.
//
// A volume is one orbital of one configuration: L^3 complex values [x][y][z], z fastest
// (the reference's site order, src/utils/vector3.C:175-187).  A 3D DFT is split into
//   * the X pass : one thread per x-pencil (y, z); the L values at stride L^2 are loaded
//                  with lanes on consecutive (y, z), so every load is a fully coalesced
//                  row of the volume, and the L-point DFT runs in registers;
//   * the YZ pass: one workgroup per PPB consecutive yz-planes; the planes are staged in
//                  LDS with one element of row padding, rows (z) and columns (y) are
//                  transformed in registers by one thread per row / column.
// Each pass reads and writes every element once.  Kinetic / field multipliers and the
// plane-wave sink measurement are applied inside the passes (the reference runs them as
// separate kernels per orbital: src/hamiltonian/cuda/hamiltonian_kernels.cu:17-57).
//
// Both passes only stream: there is no data reuse between workgroups, so blocks map to
// XCDs in launch order without remapping.
#include "nrc_ops.h"

namespace nrc {

// ------------------------------------------------------------------ X pass
template <typename T, int L>
__global__ __launch_bounds__(256) void k_xpass(PassArgs a) {
  using C = C2<T>;
  constexpr int P = L * L;
  constexpr long V = (long)L * P;
  const long gid = (long)blockIdx.x * 256 + threadIdx.x;
  if (gid >= a.nvol * P) return;
  const long vol = gid / P;
  const int yz = (int)(gid - vol * P);
  C* base = reinterpret_cast<C*>(a.psi) + vol * V + yz;
  C v[L];
#pragma unroll
  for (int x = 0; x < L; ++x) v[x] = base[(long)x * P];
  if (a.fft_a) fft_reg<T, L, +1>(v);
  const long cfg = vol / a.norb;
  apply_mul<T, L>(v, a.m1, cfg, yz, P, V);
  if (a.gather) gather_pencil<T, L>(v, a.g, yz, cfg, (int)(vol - cfg * a.norb), a.norb);
  apply_mul<T, L>(v, a.m2, cfg, yz, P, V);
  if (a.fft_b) fft_reg<T, L, -1>(v);
#pragma unroll
  for (int x = 0; x < L; ++x) base[(long)x * P] = v[x];
}

// ------------------------------------------------------------------ YZ pass
template <typename T, int L, int PPB>
__global__ __launch_bounds__(L * PPB) void k_yzpass(PassArgs a) {
  using C = C2<T>;
  constexpr int P = L * L, LP = L + 1, NT = L * PPB;
  constexpr long V = (long)L * P;
  __shared__ C sm[PPB * L * LP];
  const long nplanes = a.nvol * L;
  const long plane0 = (long)blockIdx.x * PPB;
  const int tid = threadIdx.x;
  C* g = reinterpret_cast<C*>(a.psi) + plane0 * P;
  const long left = nplanes - plane0;
  const int avail = (int)((left < PPB ? left : PPB) * P);
  for (int i = tid; i < avail; i += NT) {
    const int pl = i / P, r = i - pl * P;
    sm[pl * L * LP + (r / L) * LP + r % L] = g[i];
  }
  __syncthreads();
  const int pl = tid / L, q = tid - (tid / L) * L;
  C* s = sm + pl * L * LP;
  C v[L];
  if (a.fft_a) {  // rows along z, thread q owns row y = q
#pragma unroll
    for (int z = 0; z < L; ++z) v[z] = s[q * LP + z];
    fft_reg<T, L, -1>(v);
#pragma unroll
    for (int z = 0; z < L; ++z) s[q * LP + z] = v[z];
    __syncthreads();
  }
  // columns along y, thread q owns column z = q
#pragma unroll
  for (int y = 0; y < L; ++y) v[y] = s[y * LP + q];
  if (a.fft_a) fft_reg<T, L, -1>(v);
  const long gp = plane0 + pl;
  if (gp < nplanes) {
    const long vol = gp / L;
    const int x = (int)(gp - vol * L);
    const long cfg = vol / a.norb;
    if (a.gather) gather_pencil<T, L>(v, a.g, x * L + q, cfg, (int)(vol - cfg * a.norb), a.norb);
    apply_mul<T, L>(v, a.m1, cfg, (long)x * P + q, L, V);
  }
  if (a.fft_b) fft_reg<T, L, +1>(v);
#pragma unroll
  for (int y = 0; y < L; ++y) s[y * LP + q] = v[y];
  __syncthreads();
  if (a.fft_b) {
#pragma unroll
    for (int z = 0; z < L; ++z) v[z] = s[q * LP + z];
    fft_reg<T, L, +1>(v);
#pragma unroll
    for (int z = 0; z < L; ++z) s[q * LP + z] = v[z];
    __syncthreads();
  }
  for (int i = tid; i < avail; i += NT) {
    const int p2 = i / P, r = i - p2 * P;
    g[i] = sm[p2 * L * LP + (r / L) * LP + r % L];
  }
}

// planes per workgroup: keeps the LDS tile at or below 64 KiB
template <typename T, int L>
constexpr int yz_ppb() {
  constexpr int bytes = L * (L + 1) * (int)sizeof(C2<T>);
  constexpr int by_lds = 65536 / bytes;
  constexpr int by_thr = 256 / L;
  constexpr int p = by_lds < by_thr ? by_lds : by_thr;
  return p < 1 ? 1 : p;
}

template <typename T, int L>
static hipError_t launch_x(const PassArgs& a, hipStream_t s) {
  const long total = a.nvol * L * L;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL((k_xpass<T, L>), dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <typename T, int L>
static hipError_t launch_yz(const PassArgs& a, hipStream_t s) {
  constexpr int PPB = yz_ppb<T, L>();
  const long nplanes = a.nvol * L;
  const unsigned blocks = (unsigned)((nplanes + PPB - 1) / PPB);
  hipLaunchKernelGGL((k_yzpass<T, L, PPB>), dim3(blocks), dim3(L * PPB), 0, s, a);
  return hipGetLastError();
}

bool pass_supported(int L) {
  return L == 4 || L == 6 || L == 8 || L == 12 || L == 16 || L == 24 || L == 32;
}

#define NRC_DISPATCH_L(FN, T, a, s)            \
  switch ((a).L) {                             \
    case 4: return FN<T, 4>(a, s);             \
    case 6: return FN<T, 6>(a, s);             \
    case 8: return FN<T, 8>(a, s);             \
    case 12: return FN<T, 12>(a, s);           \
    case 16: return FN<T, 16>(a, s);           \
    case 24: return FN<T, 24>(a, s);           \
    case 32: return FN<T, 32>(a, s);           \
    default: return hipErrorInvalidValue;      \
  }

hipError_t launch_xpass(int prec, const PassArgs& a, hipStream_t s) {
  if (a.nvol <= 0) return hipSuccess;
  if (prec == NRC_PREC_F64) { NRC_DISPATCH_L(launch_x, double, a, s) }
  NRC_DISPATCH_L(launch_x, float, a, s)
}

hipError_t launch_yzpass(int prec, const PassArgs& a, hipStream_t s) {
  if (a.nvol <= 0) return hipSuccess;
  if (prec == NRC_PREC_F64) { NRC_DISPATCH_L(launch_yz, double, a, s) }
  NRC_DISPATCH_L(launch_yz, float, a, s)
}

// ------------------------------------------------------------------ element-wise kernels
template <typename T>
__global__ void k_scale_real(C2<T>* buf, const T* table, long n, long vol, int op) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const T t = table[i % vol];
    C2<T> v = buf[i];
    buf[i] = op ? C2<T>{v.x / t, v.y / t} : C2<T>{v.x * t, v.y * t};
  }
}

template <typename T>
__global__ void k_scale_const(C2<T>* buf, long n, double d, int op) {
  const T t = (T)d;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const C2<T> v = buf[i];
    buf[i] = op ? C2<T>{v.x / t, v.y / t} : C2<T>{v.x * t, v.y * t};
  }
}

// sum of squares of the 2 vol reals of volume blockIdx.x, FP64 accumulation
template <typename T>
__global__ __launch_bounds__(256) void k_norm2(const C2<T>* buf, long vol, double* out) {
  __shared__ double red[4];
  const C2<T>* b = buf + (long)blockIdx.x * vol;
  double acc = 0.0;
  for (long i = threadIdx.x; i < vol; i += 256) {
    const C2<T> v = b[i];
    acc += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
  }
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

template <typename T>
__global__ void k_mul_complex(C2<T>* buf, const C2<T>* f, long n, long vol) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const C2<T> a = f[i % vol], b = buf[i];
    buf[i] = {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
  }
}

static unsigned grid_for(long n) {
  long b = (n + 255) / 256;
  return (unsigned)(b < 65536 ? (b < 1 ? 1 : b) : 65536);
}

hipError_t launch_scale_real(int prec, void* buf, const void* table, long batch, long vol,
                             int op, hipStream_t s) {
  const long n = batch * vol;
  if (prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_scale_real<double>, dim3(grid_for(n)), dim3(256), 0, s,
                       (C2<double>*)buf, (const double*)table, n, vol, op);
  else
    hipLaunchKernelGGL(k_scale_real<float>, dim3(grid_for(n)), dim3(256), 0, s,
                       (C2<float>*)buf, (const float*)table, n, vol, op);
  return hipGetLastError();
}

hipError_t launch_scale_const(int prec, void* buf, long n, double d, int op, hipStream_t s) {
  if (prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_scale_const<double>, dim3(grid_for(n)), dim3(256), 0, s,
                       (C2<double>*)buf, n, d, op);
  else
    hipLaunchKernelGGL(k_scale_const<float>, dim3(grid_for(n)), dim3(256), 0, s,
                       (C2<float>*)buf, n, d, op);
  return hipGetLastError();
}

hipError_t launch_norm2(int prec, const void* buf, long batch, long vol, double* out,
                        hipStream_t s) {
  if (batch < 1) return hipSuccess;
  if (prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_norm2<double>, dim3((unsigned)batch), dim3(256), 0, s,
                       (const C2<double>*)buf, vol, out);
  else
    hipLaunchKernelGGL(k_norm2<float>, dim3((unsigned)batch), dim3(256), 0, s,
                       (const C2<float>*)buf, vol, out);
  return hipGetLastError();
}

hipError_t launch_mul_complex(int prec, void* buf, const void* field, long batch, long vol,
                              hipStream_t s) {
  const long n = batch * vol;
  if (prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_mul_complex<double>, dim3(grid_for(n)), dim3(256), 0, s,
                       (C2<double>*)buf, (const C2<double>*)field, n, vol);
  else
    hipLaunchKernelGGL(k_mul_complex<float>, dim3(grid_for(n)), dim3(256), 0, s,
                       (C2<float>*)buf, (const C2<float>*)field, n, vol);
  return hipGetLastError();
}

// field = (field + 1) / V: the last step of Interaction::Initialize
// (src/hamiltonian/cuda/hamiltonian_kernels.cu:59-71)
template <typename T>
__global__ void k_plus_one_div_vol(C2<T>* f, long vol) {
  const T inv = T(1) / (T)vol;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < vol;
       i += (long)gridDim.x * blockDim.x) {
    const C2<T> v = f[i];
    f[i] = {(v.x + T(1)) / (T)vol, v.y * inv};
  }
}

hipError_t launch_plus_one_div_vol(int prec, void* buf, long vol, hipStream_t s) {
  if (prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_plus_one_div_vol<double>, dim3(grid_for(vol)), dim3(256), 0, s,
                       (C2<double>*)buf, vol);
  else
    hipLaunchKernelGGL(k_plus_one_div_vol<float>, dim3(grid_for(vol)), dim3(256), 0, s,
                       (C2<float>*)buf, vol);
  return hipGetLastError();
}

__global__ void k_fill(uint32_t* p, uint32_t v4, size_t nwords) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v4;
}

__global__ void k_fill_bytes(uint8_t* p, uint8_t v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// memory clearing by kernel (no memset nodes, so the slice sequence can be graph-captured)
hipError_t launch_fill(void* p, uint8_t v, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  if (((uintptr_t)p & 3) == 0 && (bytes & 3) == 0) {
    const uint32_t v4 = 0x01010101u * v;
    hipLaunchKernelGGL(k_fill, dim3(grid_for((long)(bytes / 4))), dim3(256), 0, s,
                       (uint32_t*)p, v4, bytes / 4);
  } else {
    hipLaunchKernelGGL(k_fill_bytes, dim3(grid_for((long)bytes)), dim3(256), 0, s,
                       (uint8_t*)p, v, bytes);
  }
  return hipGetLastError();
}

// psi[v] = unit vector at site_of_orb[v % norb] (OneBody::Set(kx,ky,kz), one_body.C:215-228)
template <typename T>
__global__ void k_unit_sources(C2<T>* psi, const int* site_of_orb, long nvol, int norb,
                               long vol) {
  const long n = nvol * vol;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long v = i / vol, site = i - v * vol;
    const T one = sizeof(T) == 4 ? (T)source_amplitude(v / norb) : T(1);
    psi[i] = {site == site_of_orb[v % norb] ? one : T(0), T(0)};
  }
}

hipError_t launch_set_unit_sources(int prec, void* psi, const int* site_of_orb, long nvol,
                                   int norb, long vol, hipStream_t s) {
  const long n = nvol * vol;
  if (prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_unit_sources<double>, dim3(grid_for(n)), dim3(256), 0, s,
                       (C2<double>*)psi, site_of_orb, nvol, norb, vol);
  else
    hipLaunchKernelGGL(k_unit_sources<float>, dim3(grid_for(n)), dim3(256), 0, s,
                       (C2<float>*)psi, site_of_orb, nvol, norb, vol);
  return hipGetLastError();
}

// psi[v](x) = e^{2 pi i k.x / L} / sqrt(V), k = moms[v % norb]; the phase index is reduced
// mod L in integers and evaluated in FP64.
template <typename T>
__global__ void k_plane_waves(C2<T>* psi, const int* moms, long nvol, int norb, int L) {
  const long vol = (long)L * L * L;
  const long n = nvol * vol;
  const double nrm = 1.0 / sqrt((double)vol);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long v = i / vol, site = i - v * vol;
    const int o = (int)(v % norb);
    const int x = (int)(site / (L * L)), y = (int)((site / L) % L), z = (int)(site % L);
    const long ph = ((long)moms[3 * o] * x + (long)moms[3 * o + 1] * y + (long)moms[3 * o + 2] * z) % L;
    double sn, cs;
    sincospi(2.0 * (double)ph / (double)L, &sn, &cs);
    const double a = sizeof(T) == 4 ? nrm * source_amplitude(v / norb) : nrm;
    psi[i] = {(T)(cs * a), (T)(sn * a)};
  }
}

hipError_t launch_set_plane_waves(int prec, void* psi, const int* moms, long nvol, int norb,
                                  int L, hipStream_t s) {
  const long n = nvol * (long)L * L * L;
  if (prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_plane_waves<double>, dim3(grid_for(n)), dim3(256), 0, s,
                       (C2<double>*)psi, moms, nvol, norb, L);
  else
    hipLaunchKernelGGL(k_plane_waves<float>, dim3(grid_for(n)), dim3(256), 0, s,
                       (C2<float>*)psi, moms, nvol, norb, L);
  return hipGetLastError();
}

}  // namespace nrc
