// nrc_internal.h -- host-side declarations shared by the nrCPS HIP library sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nrcps.h"

namespace nrc {

enum MulKind { MUL_NONE = 0, MUL_TABLE = 1, MUL_FIELD = 2, MUL_RFIELD = 3 };

// A real per-site multiplier applied inside an FFT pass.
//   MUL_TABLE : table[site]                       (same for every volume)
//   MUL_FIELD : a0 + a1 * phi(site), phi = +-1 from the configuration's packed bits; with a
//               table (a trap) table[site] for phi = +1 and table[V + site] for phi = -1,
//               the factors (a0 +- a1) q(site) in MUL_TABLE format
//   MUL_RFIELD: table[config * V + site]          (per-configuration real field)
//
// In FP32 passes a MUL_TABLE entry is an interleaved (hi, lo) float pair applied as
// v*hi + v*lo, with hi the float nearest to the FP64 value and lo the float nearest to the
// remainder, so the factor carries ~48 bits and its rounding does not bias the evolution
// slice after slice; FP64 passes hold one double per site.  MUL_FIELD factors are split the
// same way in registers.
struct MulSpec {
  int kind = MUL_NONE;
  const void* table = nullptr;     // device: MUL_TABLE entries, or the MUL_RFIELD field
  const uint32_t* bits = nullptr;  // device, [config][nw]
  int nw = 0;
  double a0 = 0.0, a1 = 0.0;
};

// Plane-wave sink measurement folded into a pass: for each pencil p of a volume, the
// entries ptr[p] .. ptr[p+1] name (position within the pencil, sink index); the value at
// that position (times `scale`) is written to M[(config * nsink + sink) * norb + orbital].
//
// Conjugate partners (col != null): volume `orbital` of a configuration holds orbital
// col[orbital] and M has ncol columns.  When pcol[orbital] >= 0 the volume also stands for
// the orbital of opposite momentum, whose state is the conjugate mirror of this one
// (psi_{-k}(p) = conj psi_k(-p) for real fields and even kinetic tables); the entries
// ptr2 / pos2 / sink2 name the sites -k_s, and conj(value) goes to column pcol[orbital].
struct GatherSpec {
  const int* ptr = nullptr;
  const int* pos = nullptr;
  const int* sink = nullptr;
  double2* M = nullptr;
  int nsink = 0;
  double scale = 1.0;
  const double* sink_scale = nullptr;  // per-sink factor (null: 1)
  const int* col = nullptr;
  const int* pcol = nullptr;
  const int* ptr2 = nullptr;
  const int* pos2 = nullptr;
  const int* sink2 = nullptr;
  int ncol = 0;
};

struct PassArgs {
  void* psi = nullptr;  // volumes [nvol][L][L][L] complex (float2 / double2)
  long nvol = 0;        // configurations * orbitals
  int norb = 1;         // orbitals per configuration (volume v belongs to config v / norb)
  int L = 0;
  // X pass: fft_a = sign +1 x-transform first, fft_b = sign -1 x-transform last.
  // YZ pass: fft_a = sign -1 yz-transform first, fft_b = sign +1 yz-transform last.
  int fft_a = 0, fft_b = 0;
  MulSpec m1, m2;  // X pass: m1 before the gather, m2 after.  YZ pass: m1 between the FFTs.
  int gather = 0;
  GatherSpec g;
};

hipError_t launch_xpass(int prec, const PassArgs& a, hipStream_t s);
hipError_t launch_yzpass(int prec, const PassArgs& a, hipStream_t s);
bool pass_supported(int L);
// 3D DFTs of X x Y x Z lattices, one launch per axis (fft_axis.hip): sides in {2, 4, 6, 8, 12,
// 16, 24, 32, 48, 64}
bool axis_supported(int n);
hipError_t launch_fft3d_axes(int prec, void* buf, int lx, int ly, int lz, long batch, int sign,
                             hipStream_t st);

// One whole slice per orbital volume in a single HBM pass (slice_fused.hip):
// pre multiplier, x/y/z DFT (-1), mid multiplier, DFT (+1), post multiplier, with the sink
// gather at the start (gather_pre, x-pencils) or between the transforms (gather_mid,
// (x, z) columns).
struct SliceArgs {
  void* psi = nullptr;
  long nvol = 0;
  int norb = 1;
  int L = 0;
  MulSpec pre, mid, post;
  int gather_pre = 0, gather_mid = 0;
  GatherSpec g;
  int grid = 0;  // persistent kernel: the device's CU count (0: one workgroup per volume)
  // persistent kernel, k-space split: nslices consecutive slices per launch; slice s > 0
  // starts with pre_next, reads its field words at mid.bits + s * bits_stride and the sinks
  // gathered after slice s go to g.M + s * m_stride
  int nslices = 1;
  const void* pre_next = nullptr;
  long bits_stride = 0;
  long m_stride = 0;
};
constexpr int SLICE_FUSE_MAX = 4;  // slices per launch of the persistent slice kernel

// The k-space split slice loop with FP32 state and the x-chain in FP64 (slice_ext.hip):
// `nslices` slices per launch of the persistent kernel.  The volumes hold a k-space state
// (start_k: the sources when half0, else the folded state) or the state of the previous
// launch in (x, k_y, k_z) coordinates with its last x-DFT(+1) pending; slice s of the launch
// gathers its measurement (the sinks of g, g.M advanced by s * m_stride; sink_scale except
// on the sources), multiplies by the FP64 table tab_x = K(1/2) (sources) or tab_k = K, and
// evolves with the field words at bits + s * bits_stride: (a0 + a1 phi) in x-space, times
// the separable trap when trap_sep is set.  store_k: finish with the x-DFT(+1) and store the
// folded k-space state.
struct SliceXArgs {
  void* psi = nullptr;  // float2 volumes
  long nvol = 0;
  int norb = 1;
  int L = 0;
  int grid = 0;  // CUs
  int nslices = 1;
  int start_k = 0, store_k = 0;
  int half0 = 0;   // slice 0 acts on the sources: factor K(1/2), gather without sink_scale
  int gather = 1;  // gather the plane-wave sinks of g (0: measured elsewhere)
  const double* tab_x = nullptr;
  const double* tab_k = nullptr;
  const uint32_t* bits = nullptr;
  int nw = 0;
  long bits_stride = 0;
  double a0 = 0.0, a1 = 0.0;
  const double* trap_sep = nullptr;  // separable trap q = qx qy qz: [qx(L), qy(L), qz(L)]
  GatherSpec g;
  long m_stride = 0;
  int tile = 0;  // x-pencils per thread (0: the shape's default; slicex_tiles)
  int zd = 1;    // z rows in FP64 where the shape allows it (0: FP32 z rows)
  // with nslices > 1: the folded k-space state after each slice but the last, FP32 volumes
  // at snap + k snap_stride (k = 0 .. nslices - 2, snap_stride complex elements apart)
  void* snap = nullptr;
  long snap_stride = 0;
};
bool slicex_supported(int L);
int slicex_tiles(int L, int* tiles);  // the tiles a shape offers, default first; returns count
hipError_t launch_slicex(const SliceXArgs& a, hipStream_t s);
bool slice1_supported(int prec, int L);
bool slice1p_trap_ok(int prec, int L);  // the persistent kernel takes a trap at this shape
hipError_t launch_slice1(int prec, const SliceArgs& a, hipStream_t s);
// Columns of conjugate partners for the gather from memory (all null: column n is volume n):
// column n is volume vol[n] of its configuration (nvo volumes each), read at the mirrored
// sink site msite[s] and conjugated when conj[n] != 0.
struct ConjCols {
  const int* vol = nullptr;
  const int* conj = nullptr;
  const int* msite = nullptr;
  int nvo = 0;
};
// M[c][s][n] = psi[c][n][site_of_sink[s]] (plane-wave sinks on k-space orbitals)
hipError_t launch_gather_sites(int prec, const void* psi, const int* site_of_sink,
                               const double* sink_scale, double2* M, long C, int N, long V,
                               hipStream_t s, const ConjCols& cp = ConjCols());

// element-wise helpers (the reference's separate kernels, used by the class layer)
hipError_t launch_scale_real(int prec, void* buf, const void* table, long batch, long vol,
                             int op, hipStream_t s);
hipError_t launch_mul_complex(int prec, void* buf, const void* field, long batch, long vol,
                              hipStream_t s);
// buf[i] *= d (op 0) or /= d (op 1) for n complex values
hipError_t launch_scale_const(int prec, void* buf, long n, double d, int op, hipStream_t s);
// out[b] = sum over volume b of |v|^2 (FP64), device output
hipError_t launch_norm2(int prec, const void* buf, long batch, long vol, double* out,
                        hipStream_t s);
hipError_t launch_fill(void* p, uint8_t v, size_t bytes, hipStream_t s);
hipError_t launch_plus_one_div_vol(int prec, void* buf, long vol, hipStream_t s);
// sources of nvol = configurations x norb volumes; FP32 sources of configuration c carry the
// amplitude source_amplitude(c)
hipError_t launch_set_unit_sources(int prec, void* psi, const int* site_of_orb, long nvol,
                                   int norb, long vol, hipStream_t s);
hipError_t launch_set_plane_waves(int prec, void* psi, const int* moms, long nvol, int norb,
                                  int L, hipStream_t s);

// Philox Z2 field bits: words[c][w] for configs config0 .. config0 + nconf - 1
// (par != null: seed and config0 read from par[0], par[1] in device memory at run time)
hipError_t launch_z2_bits(uint64_t seed, uint32_t config0, int nconf, uint32_t t,
                          uint32_t stream, int vol, uint32_t* words, hipStream_t s,
                          const uint64_t* par = nullptr);
// Philox stream of the tilted sampler's per-site uniforms (the plain sampler uses stream 0,
// the learned proposal stream 1)
#define NRC_STREAM_TILT 2u
// Tilted Philox Z2 field bits: P(+1) = thr / 2^32 per site, one uniform per site; fsum[c]
// receives each configuration's field sum
hipError_t launch_z2_tilt(uint64_t seed, uint32_t config0, int nconf, uint32_t t,
                          uint32_t stream, int vol, uint32_t thr, uint32_t* words, int* fsum,
                          hipStream_t s, const uint64_t* par = nullptr);
// Philox stream of the memory sampler's per-site uniforms
#define NRC_STREAM_MEMORY 3u
// Philox stream of the memory network's per-site uniforms
#define NRC_STREAM_MEMNET 4u
// The memory network (proposal.hip): 35 FP32 parameters prm (device); the words of slice t,
// the per-site memories state (one word per site, updated in place; not read at t = 0) and
// lr[c], the slice's log(p / q)
hipError_t launch_z2_memnet(uint64_t seed, uint32_t config0, int nconf, uint32_t t, int vol,
                            const float* prm, uint32_t* words, uint32_t* state, double* lr,
                            hipStream_t s, const uint64_t* par = nullptr);
// The memory network on the plain sampler's fields: grad false, out[t][c] = the cumulative
// log(p / q) for t < T; grad true, out[c][35] = the gradient of sum_t G[t][c] sum log q
// (vol <= 32768)
hipError_t launch_z2_memnet_walk(uint64_t seed, uint32_t config0, int nconf, int T, int vol,
                                 const float* prm, const double* G, double* out, bool grad,
                                 hipStream_t s);
// Z2 fields with memory (integer parameters a, b, r; see lu_rng.hip): the words of slice t,
// the per-site memory mem (int16, 32 per word, updated in place; not read at t = 0) and
// lr[c], the slice's log(p / q)
hipError_t launch_z2_memory(uint64_t seed, uint32_t config0, int nconf, uint32_t t,
                            uint32_t stream, int vol, int a, int b, int r, uint32_t* words,
                            int16_t* mem, double* lr, hipStream_t s,
                            const uint64_t* par = nullptr);
// The memory proposal's cumulative log(p / q) over the plain sampler's fields: out[t][c] for
// t < T (vol <= 32768)
hipError_t launch_z2_memory_eval(uint64_t seed, uint32_t config0, int nconf, int T, int vol,
                                 int a, int b, int r, double* out, hipStream_t s);

// Options of the batched LU (launch_lu_det), all off by default:
//   repl, ldr: the last column of every minor (size k) is row k - 1 of repl[c] (a matrix of
//     ldr columns per row, n rows per configuration) read as a column, i.e. SlaterDet3::
//     Run(k, -1, k - 1) with psi1[s + (k-1) N] (src/slater_det/slater_det3.C:163-167);
//   lsrc, lsrc_cols: log|det| of a k x k minor of configuration c is reduced by
//     (lsrc_cols k - (repl ? lsrc_cols - 1 : 0)) lsrc[c]: lsrc_cols source amplitudes per
//     column, one in a replaced column;
//   onm, ostep, oofs: the result of minor m of configuration c goes to
//     c onm + m ostep + oofs (onm 0: the number of minors).
struct LuOpts {
  const double2* repl = nullptr;
  int ldr = 0;
  const double* lsrc = nullptr;
  int lsrc_cols = 1;
  long lsrc_period = 0;  // > 0: matrix c uses lsrc[c % lsrc_period] (several slices at once)
  int onm = 0, ostep = 1, oofs = 0;
};

// batched LU log-determinant of M[c][n][n] (leading minors when minors != 0)
hipError_t launch_lu_det(const double2* M, int n, int ldm, int ncfg, int minors,
                         double* logabs, double2* phase, double* argsum, int* status,
                         double2* lu_out, int* indx_out, double* d_out, hipStream_t s,
                         const LuOpts& opts = LuOpts());

// Amplitude r_c in [1, 2) of configuration c's FP32 sources: 23 hashed mantissa bits, so r_c
// is exact in FP32.  The first slice acts on sources that are the same in every
// configuration, and FP32 rounds their products identically everywhere, which biases
// log|det| by a constant (~4e-7 at L = 32, N = 33); scaling configuration c by r_c
// decorrelates those roundings, and since the evolution is linear the measurements divide
// out exactly (det of a k x k minor scales by r_c^k).
__host__ __device__ inline double source_amplitude(long c) {
  uint32_t h = (uint32_t)c * 0x9E3779B1u;
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return 1.0 + (double)(h >> 9) * (1.0 / 8388608.0);
}
// |z| computed exactly as the host C library's hypot (the LU's pivot metric)
hipError_t launch_ref_abs(const double2* z, double* out, long n, hipStream_t s);

// dense overlaps M[c][s][t] = sum_i conj(phi[s][i]) psi[c][t][i]; the overlap entry points
// take configuration c's orbitals at psi + c cstride (complex elements; 0: nt vol), so a call
// can cover a range of the volumes a configuration stores
hipError_t launch_overlap_dense(int prec, const void* phi, const void* psi, double2* M,
                                int ns, int nt, long vol, int ncfg, hipStream_t s,
                                long cstride = 0);
// terms 3: split-BF16x3 (phi_split: hi, lo planes); 6: split-BF16x6 (three planes)
hipError_t launch_overlap_bf16x3(const void* phi_split, const float2* psi, double2* M, int ns,
                                 int nt, long vol, int ncfg, float* partial, hipStream_t s,
                                 int terms = 3, long cstride = 0);
// phi_split: `parts` bf16 planes [part][ns][2V] (hi, lo, and with parts = 3 the third part)
hipError_t launch_split_bf16(const float2* phi, void* out, int ns, long vol, hipStream_t s,
                             int parts = 2);
size_t overlap_bf16x3_partial_floats(int ns, int nt, long vol, int ncfg);
// the sink-row form: rows_split holds `parts` planes [part][2 ns][2V] (launch_split_rows),
// row 2s = (re, im) and row 2s + 1 = (-im, re) of sink s, staged in LDS and shared by the four
// waves of a workgroup, each a tile of orbital columns (c, t) taken across configurations;
// form 0: 2-step stages, 2 steps of orbital loads in flight,
// two workgroups per CU; 1: 4-step stages, 4 in flight, one per CU; 2: 2-step stages, 4 in
// flight, two per CU
hipError_t launch_split_rows(const float2* phi, void* out, int ns, long vol, hipStream_t s,
                             int parts);
hipError_t launch_overlap_lds(const void* rows_split, const float2* psi, double2* M, int ns,
                              int nt, long vol, int ncfg, float* partial, hipStream_t s,
                              int terms, int form, long cstride = 0);
size_t overlap_rows_partial_floats(int ns, int nt, long vol, int ncfg);

// pair-sink matrices M[c][s][t] = sum_i w[i] psi1[c][s][i] psi2[c][t][opp[i]]: FP64 sums
// (dense), or split-BF16x3 on the matrix cores (FP32 orbitals, weights rounded to FP32).
// opp null: conj(psi2[c][t][i]) at the same site (a conjugate partner's mirror image).
hipError_t launch_pair_dense(int prec, const void* psi1, const void* psi2, const double* w,
                             const int* opp, double2* M, int n1, int n2, long vol, int ncfg,
                             hipStream_t s);
hipError_t launch_pair_bf16x3(const float2* psi1, const float2* psi2, const double* wf,
                              const int* opp, double2* M, int n1, int n2, long vol, int ncfg,
                              float* partial, hipStream_t s);
size_t pair_bf16x3_partial_floats(int n1, int n2, long vol, int ncfg);
// G[c][s][t] = sum_i w[i] psi[c][s][i] psi[c][t][opp[i]] and H[c][s][t] = sum_i w[i]
// psi[c][s][i] conj psi[c][t][i] in one pass (split-BF16x3 on the matrix cores; n <= 48).
// L > 0: opp is the mirror p -> -p of a cubic lattice of side L (a multiple of 4, >= 8) and w
// is even; each site and its mirror are then loaded together (opp and wf unused) and ws[V]
// holds sg(w) sqrt|w| in FP32 (ws_neg: some w is negative)
hipError_t launch_pair_gh_bf16x3(const float2* psi, const double* wf, const int* opp,
                                 double2* G, double2* H, int n, long vol, int ncfg,
                                 float* partial, hipStream_t s, int L = 0,
                                 const float* ws = nullptr, bool ws_neg = false);
size_t pair_gh_bf16x3_partial_floats(int n, long vol, int ncfg, int L = 0);
// the N x N pair matrices of configurations whose orbitals are stored as nvo volumes with
// conjugate partners, from G[c][a][b] = sum w psi_a(p) psi_b(-p) and H[c][a][b] =
// sum w psi_a(p) conj psi_b(p) over the stored volumes: orbital s is volume col_vol[s],
// mirrored when col_conj[s] (M = G, H, conj H, conj G for the four cases)
// the N x N overlaps of configurations stored as nvo volumes with conjugate partners, from
// the overlaps of the sinks with the stored volumes (Mp) and of their conjugate mirrors with
// the volumes v0 .. v0 + nm - 1 (Mm)
hipError_t launch_sink_expand(const double2* Mp, const double2* Mm, const int* col_vol,
                              const int* col_conj, double2* M, int N, int nvo, int nm, int v0,
                              long ncfg, hipStream_t s);
hipError_t launch_pair_expand(const double2* G, const double2* H, const int* col_vol,
                              const int* col_conj, double2* M, int N, int nvo, long ncfg,
                              hipStream_t s);

// the context's stream and device (api.hip), for entry points defined in other files
hipStream_t ctx_stream(nrc_context* c);
int ctx_device(nrc_context* c);
int ctx_n_cu(nrc_context* c);

}  // namespace nrc
