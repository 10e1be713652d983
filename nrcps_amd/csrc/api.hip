// api.hip -- the extern "C" entry points of libnrcps_hip.so (declared in include/nrcps.h).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <map>
#include <vector>

#include "nrc_device.h"
#include "nrc_internal.h"

struct nrc_context {
  int device = 0;
  hipStream_t stream = nullptr;
  char name[256] = {0};
  int n_cu = 0;
  void* work = nullptr;  // grow-only workspace of the library's own calls
  size_t work_bytes = 0;
  // tuned choices (nrc_overlap_tune, nrc_ensemble_tune; nrc_tune_save / nrc_tune_load):
  // GEMM variant per (ns, nt, vol, terms); (path, slices per launch) per ensemble shape key
  std::map<std::array<long, 4>, int> ovl_choice;
  std::map<std::array<long, 8>, std::array<int, 3>> ens_choice;  // path, slices, tile
};

// The scratch of the no-allocation overlap (nrc_overlap_noalloc): a fixed device array that
// every GPU the library runs on holds until the process ends.  Host code never uses the
// array's name as an address (on the host it is a stand-in symbol): nrc_init looks up each
// device's own copy with hipGetSymbolAddress, with that device current, and keeps the address
// in g_scratch_dev[device]; that address is what kernels and copies get.
constexpr size_t NRC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NRC_SCRATCH_BYTES];
constexpr int NRC_MAX_DEVICES = 64;
static void* g_scratch_dev[NRC_MAX_DEVICES] = {nullptr};

// the context workspace, grown to at least `bytes` (synchronises before a reallocation)
static void* ctx_workspace(nrc_context* c, size_t bytes) {
  if (bytes <= c->work_bytes) return c->work;
  hipStreamSynchronize(c->stream);
  if (c->work) hipFree(c->work);
  c->work = nullptr;
  c->work_bytes = 0;
  if (hipMalloc(&c->work, bytes) != hipSuccess) return nullptr;
  c->work_bytes = bytes;
  return c->work;
}

namespace nrc {
hipStream_t ctx_stream(nrc_context* c) { return c->stream; }
int ctx_device(nrc_context* c) { return c->device; }
int ctx_n_cu(nrc_context* c) { return c->n_cu; }
}  // namespace nrc

struct nrc_fft {
  nrc_context* ctx;
  int L, batch, prec;
  int lx, ly, lz;  // the lattice; fused = cubic L <= 32 on the fused passes, else per axis
  bool fused;
};

// ------------------------------------------------------------------ failures and check mode
// Every failure records "file:line: what" for its thread (nrc_last_error_detail).  A build
// with -DNRC_CHECK (libnrcps_hip_check.so) also synchronises the device after every HIP call
// outside graph capture, so that an asynchronous kernel failure is reported by the entry
// point that launched it; checks the contents of index arrays a caller hands to a kernel
// before the kernel runs; and scans results for NaN (and infinities where none may occur),
// returning NRC_ERR_CHECK.  The default build does none of the extra work.
static thread_local char g_detail[256] = "";
static thread_local int g_capturing = 0;  // a graph capture is open on this thread

static void set_detail(const char* file, int line, const char* what) {
  const char* base = strrchr(file, '/');
  snprintf(g_detail, sizeof(g_detail), "%s:%d: %s", base ? base + 1 : file, line, what);
}

#ifdef NRC_CHECK
#define NRC_SYNC_CHECK(e_)                                       \
  if (e_ == hipSuccess && !g_capturing) {                        \
    e_ = hipDeviceSynchronize();                                 \
    if (e_ == hipSuccess) e_ = hipGetLastError();                \
  }
#else
#define NRC_SYNC_CHECK(e_)
#endif

#define NRC_HIP(expr)                                            \
  do {                                                           \
    hipError_t e_ = (expr);                                      \
    NRC_SYNC_CHECK(e_)                                           \
    if (e_ != hipSuccess) {                                      \
      set_detail(__FILE__, __LINE__, hipGetErrorString(e_));     \
      return NRC_ERR_HIP;                                        \
    }                                                            \
  } while (0)

#define NRC_TRY(expr)                                  \
  do {                                                 \
    int r_ = (expr);                                   \
    if (r_ != NRC_OK) return r_;                       \
  } while (0)

#ifdef NRC_CHECK
// count of values that are NaN, or infinite unless allow_neg_inf and negative
template <typename T>
__global__ void k_check_finite(const T* p, long n, int allow_neg_inf, int* bad) {
  int local = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const double x = (double)p[i];
    if (isnan(x) || (isinf(x) && !(allow_neg_inf && x < 0.0))) ++local;
  }
  if (local) atomicAdd(bad, local);
}

// count of indices outside [lo, hi)
__global__ void k_check_range(const int* p, long n, int lo, int hi, int* bad) {
  int local = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    if (p[i] < lo || p[i] >= hi) ++local;
  if (local) atomicAdd(bad, local);
}

static int check_count(hipStream_t s, int* bad_dev, const char* what, const char* file,
                       int line) {
  int bad = 0;
  NRC_HIP(hipMemcpyAsync(&bad, bad_dev, sizeof(int), hipMemcpyDeviceToHost, s));
  NRC_HIP(hipStreamSynchronize(s));
  hipFree(bad_dev);
  if (!bad) return NRC_OK;
  char msg[160];
  snprintf(msg, sizeof(msg), "%s: %d bad values", what, bad);
  set_detail(file, line, msg);
  return NRC_ERR_CHECK;
}

template <typename T>
static int check_finite(hipStream_t s, const T* p, long n, bool allow_neg_inf, const char* what,
                        const char* file, int line) {
  if (g_capturing || n <= 0) return NRC_OK;
  int* bad = nullptr;
  NRC_HIP(hipMalloc(&bad, sizeof(int)));
  NRC_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  const long blocks = std::min<long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_check_finite<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, n,
                     (int)allow_neg_inf, bad);
  return check_count(s, bad, what, file, line);
}

static int check_range(hipStream_t s, const int* p, long n, int lo, int hi, const char* what,
                       const char* file, int line) {
  if (g_capturing || n <= 0) return NRC_OK;
  int* bad = nullptr;
  NRC_HIP(hipMalloc(&bad, sizeof(int)));
  NRC_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  const long blocks = std::min<long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_check_range, dim3((unsigned)blocks), dim3(256), 0, s, p, n, lo, hi, bad);
  return check_count(s, bad, what, file, line);
}
#define NRC_CHECK_FINITE(s, p, n, allow_neg_inf, what) \
  NRC_TRY(check_finite(s, p, n, allow_neg_inf, what, __FILE__, __LINE__))
#define NRC_CHECK_RANGE(s, p, n, lo, hi, what) \
  NRC_TRY(check_range(s, p, n, lo, hi, what, __FILE__, __LINE__))
#else
#define NRC_CHECK_FINITE(s, p, n, allow_neg_inf, what) \
  do {                                                \
  } while (0)
#define NRC_CHECK_RANGE(s, p, n, lo, hi, what) \
  do {                                        \
  } while (0)
#endif

static size_t elem_bytes(int prec) { return prec == NRC_PREC_F64 ? 16 : 8; }
static bool prec_ok(int prec) { return prec == NRC_PREC_F32 || prec == NRC_PREC_F64; }

extern "C" {

const char* nrc_error_string(int code) {
  switch (code) {
    case NRC_OK: return "ok";
    case NRC_ERR_ARG: return "invalid argument";
    case NRC_ERR_UNSUPPORTED: return "unsupported shape or mode";
    case NRC_ERR_HIP: return "HIP runtime error";
    case NRC_ERR_NOMEM: return "device allocation failed";
    case NRC_ERR_SINGULAR: return "singular matrix (all-zero row)";
    case NRC_ERR_BATCHED: return "ensemble belongs to a batch";
    case NRC_ERR_TOO_LARGE: return "sizes exceed the fixed scratch of the no-allocation call";
    case NRC_ERR_CHECK: return "a check of the NRC_CHECK build failed";
    default: return "unknown error";
  }
}

const char* nrc_last_error_detail(void) { return g_detail; }

int nrc_check_build(void) {
#ifdef NRC_CHECK
  return 1;
#else
  return 0;
#endif
}

int nrc_init(int device, nrc_context** ctx) {
  if (!ctx) return NRC_ERR_ARG;
  int n = 0;
  NRC_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return NRC_ERR_ARG;
  NRC_HIP(hipSetDevice(device));
  nrc_context* c = new nrc_context();
  c->device = device;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) { delete c; return NRC_ERR_HIP; }
  strncpy(c->name, p.gcnArchName, sizeof(c->name) - 1);
  c->n_cu = p.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return NRC_ERR_HIP;
  }
  if (device < NRC_MAX_DEVICES && !g_scratch_dev[device]) {  // once per device, device current
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
      hipStreamDestroy(c->stream);
      delete c;
      return NRC_ERR_HIP;
    }
    g_scratch_dev[device] = ptr;
  }
  *ctx = c;
  return NRC_OK;
}

int nrc_finalize(nrc_context* ctx) {
  if (!ctx) return NRC_ERR_ARG;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  if (ctx->work) hipFree(ctx->work);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return NRC_OK;
}

int nrc_device_name(nrc_context* ctx, char* buf, int len, int* n_cu) {
  if (!ctx || !buf || len < 1) return NRC_ERR_ARG;
  strncpy(buf, ctx->name, (size_t)len - 1);
  buf[len - 1] = 0;
  if (n_cu) *n_cu = ctx->n_cu;
  return NRC_OK;
}

int nrc_malloc(nrc_context* ctx, void** p, size_t bytes) {
  if (!ctx || !p) return NRC_ERR_ARG;
  hipSetDevice(ctx->device);
  if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) return NRC_ERR_NOMEM;
  return NRC_OK;
}

int nrc_free(nrc_context* ctx, void* p) {
  if (!ctx) return NRC_ERR_ARG;
  if (!p) return NRC_OK;
  NRC_HIP(hipStreamSynchronize(ctx->stream));
  NRC_HIP(hipFree(p));
  return NRC_OK;
}

int nrc_memcpy_h2d(nrc_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return NRC_ERR_ARG;
  NRC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  NRC_HIP(hipStreamSynchronize(ctx->stream));
  return NRC_OK;
}

int nrc_memcpy_d2h(nrc_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return NRC_ERR_ARG;
  NRC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  NRC_HIP(hipStreamSynchronize(ctx->stream));
  return NRC_OK;
}

int nrc_memcpy_d2d(nrc_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return NRC_ERR_ARG;
  NRC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return NRC_OK;
}

int nrc_memset(nrc_context* ctx, void* p, int byte, size_t bytes) {
  if (!ctx || (!p && bytes)) return NRC_ERR_ARG;
  NRC_HIP(nrc::launch_fill(p, (uint8_t)byte, bytes, ctx->stream));
  return NRC_OK;
}

int nrc_sync(nrc_context* ctx) {
  if (!ctx) return NRC_ERR_ARG;
  NRC_HIP(hipStreamSynchronize(ctx->stream));
  return NRC_OK;
}

// ------------------------------------------------------------------ FFT
int nrc_fft_plan(nrc_context* ctx, int lx, int ly, int lz, int batch, int prec, nrc_fft** plan) {
  if (!ctx || !plan || batch < 0 || !prec_ok(prec)) return NRC_ERR_ARG;
  const bool fused = lx == ly && ly == lz && nrc::pass_supported(lx);
  if (!fused && !(nrc::axis_supported(lx) && nrc::axis_supported(ly) && nrc::axis_supported(lz)))
    return NRC_ERR_UNSUPPORTED;
  *plan = new nrc_fft{ctx, lx, batch, prec, lx, ly, lz, fused};
  return NRC_OK;
}

int nrc_fft_exec(nrc_fft* p, void* buf, int sign) {
  if (!p || !buf || (sign != -1 && sign != 1)) return NRC_ERR_ARG;
  const long nreal = 2L * p->batch * p->lx * p->ly * p->lz;
  if (!p->fused) {
    NRC_HIP(nrc::launch_fft3d_axes(p->prec, buf, p->lx, p->ly, p->lz, p->batch, sign,
                                   p->ctx->stream));
    if (p->prec == NRC_PREC_F64)
      NRC_CHECK_FINITE(p->ctx->stream, (const double*)buf, nreal, false, "fft output");
    else
      NRC_CHECK_FINITE(p->ctx->stream, (const float*)buf, nreal, false, "fft output");
    return NRC_OK;
  }
  // X pass: fft_a is the sign +1 x-transform, fft_b the sign -1 one; YZ pass: fft_a is the
  // sign -1 yz-transform, fft_b the sign +1 one.
  nrc::PassArgs x, yz;
  x.psi = yz.psi = buf;
  x.nvol = yz.nvol = p->batch;
  x.L = yz.L = p->L;
  if (sign < 0) { x.fft_b = 1; yz.fft_a = 1; } else { x.fft_a = 1; yz.fft_b = 1; }
  NRC_HIP(nrc::launch_xpass(p->prec, x, p->ctx->stream));
  NRC_HIP(nrc::launch_yzpass(p->prec, yz, p->ctx->stream));
  if (p->prec == NRC_PREC_F64)
    NRC_CHECK_FINITE(p->ctx->stream, (const double*)buf, nreal, false, "fft output");
  else
    NRC_CHECK_FINITE(p->ctx->stream, (const float*)buf, nreal, false, "fft output");
  return NRC_OK;
}

int nrc_fft_destroy(nrc_fft* p) {
  delete p;
  return NRC_OK;
}

// ------------------------------------------------------------------ element-wise
int nrc_scale_real(nrc_context* ctx, void* buf, const void* table, int batch, long vol, int op,
                   int prec) {
  if (!ctx || !buf || !table || batch < 0 || vol < 1 || (op != 0 && op != 1) || !prec_ok(prec))
    return NRC_ERR_ARG;
  NRC_HIP(nrc::launch_scale_real(prec, buf, table, batch, vol, op, ctx->stream));
  return NRC_OK;
}

int nrc_scale_const(nrc_context* ctx, void* buf, long n, double d, int op, int prec) {
  if (!ctx || !buf || n < 0 || (op != 0 && op != 1) || !prec_ok(prec)) return NRC_ERR_ARG;
  if (op == 1 && d == 0.0) return NRC_ERR_ARG;
  if (n == 0) return NRC_OK;
  NRC_HIP(nrc::launch_scale_const(prec, buf, n, d, op, ctx->stream));
  return NRC_OK;
}

int nrc_norms(nrc_context* ctx, const void* buf, int batch, long vol, int prec, double* norms) {
  if (!ctx || !buf || !norms || batch < 0 || vol < 1 || !prec_ok(prec)) return NRC_ERR_ARG;
  if (batch == 0) return NRC_OK;
  double* d = nullptr;
  if (hipMalloc(&d, (size_t)batch * sizeof(double)) != hipSuccess) return NRC_ERR_NOMEM;
  int rc = nrc::launch_norm2(prec, buf, batch, vol, d, ctx->stream) == hipSuccess ? NRC_OK
                                                                                 : NRC_ERR_HIP;
  if (rc == NRC_OK &&
      hipMemcpyAsync(norms, d, (size_t)batch * sizeof(double), hipMemcpyDeviceToHost,
                     ctx->stream) != hipSuccess)
    rc = NRC_ERR_HIP;
  if (rc == NRC_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = NRC_ERR_HIP;
  hipFree(d);
  for (int b = 0; rc == NRC_OK && b < batch; ++b) norms[b] = sqrt(norms[b]);
  return rc;
}

int nrc_mul_complex(nrc_context* ctx, void* buf, const void* field, int batch, long vol,
                    int prec) {
  if (!ctx || !buf || !field || batch < 0 || vol < 1 || !prec_ok(prec)) return NRC_ERR_ARG;
  NRC_HIP(nrc::launch_mul_complex(prec, buf, field, batch, vol, ctx->stream));
  return NRC_OK;
}

int nrc_plus_one_div_vol(nrc_context* ctx, void* buf, long vol, int prec) {
  if (!ctx || !buf || vol < 1 || !prec_ok(prec)) return NRC_ERR_ARG;
  NRC_HIP(nrc::launch_plus_one_div_vol(prec, buf, vol, ctx->stream));
  return NRC_OK;
}

// ------------------------------------------------------------------ overlaps, LU, fields
}  // extern "C"

// The split-BF16 overlap GEMM comes in variants (NRC_OVL_*): the column form (the orbital
// side builds a real and an imaginary column per orbital, four waves per configuration), and
// the sink-row form (two rows per sink, split once, staged in LDS for the four column tiles of
// a workgroup) with 2-step or 4-step stages.  A call's variant is the context's tuned
// choice for (ns, nt, vol, terms), else the default below.
static const int kOvlVariants[] = {NRC_OVL_COLS, NRC_OVL_SHARED, NRC_OVL_SHARED_DEEP,
                                   NRC_OVL_SHARED_PIPE};
constexpr int kNumOvlVariants = 4;

static int overlap_variant(const nrc_context* c, int ns, int nt, long vol, int terms) {
  if (c) {
    const auto it = c->ovl_choice.find({ns, nt, vol, terms});
    if (it != c->ovl_choice.end()) return it->second;
  }
  return NRC_OVL_SHARED;
}

// workspace of one split-BF16 overlap call: the split sinks (up to 3 bf16 planes; 2 ns rows
// in the row form) and the per-chunk FP32 partial tiles; variant NRC_OVL_AUTO sizes for any
static size_t overlap_split_bytes(int ns, long vol, int variant) {
  const size_t rows = variant == NRC_OVL_COLS ? (size_t)ns : (size_t)2 * ns;
  return ((size_t)3 * rows * 2 * vol * 2 + 255) / 256 * 256;
}
static size_t overlap_bf16x3_workspace(int ns, int nt, long vol, int ncfg,
                                       int variant = NRC_OVL_AUTO) {
  if (variant == NRC_OVL_AUTO) {
    size_t m = 0;
    for (int v : kOvlVariants) m = std::max(m, overlap_bf16x3_workspace(ns, nt, vol, ncfg, v));
    return m;
  }
  const size_t partial = variant == NRC_OVL_COLS
                             ? nrc::overlap_bf16x3_partial_floats(ns, nt, vol, ncfg)
                             : nrc::overlap_rows_partial_floats(ns, nt, vol, ncfg);
  return overlap_split_bytes(ns, vol, variant) + partial * sizeof(float);
}
// terms 3: split-BF16x3; 6: split-BF16x6 (three-way splits, six products)
static int overlap_bf16x3(const void* phi, const void* psi, double* M, int ns, int nt, long vol,
                          int ncfg, void* work, hipStream_t s, int terms, int variant,
                          long cstride = 0) {
  char* w = (char*)work;
  const size_t split = overlap_split_bytes(ns, vol, variant);
  const int parts = terms == 6 ? 3 : 2;
  if (variant == NRC_OVL_COLS) {
    NRC_HIP(nrc::launch_split_bf16((const float2*)phi, w, ns, vol, s, parts));
    NRC_HIP(nrc::launch_overlap_bf16x3(w, (const float2*)psi, (double2*)M, ns, nt, vol, ncfg,
                                       (float*)(w + split), s, terms, cstride));
    return NRC_OK;
  }
  const int form = variant == NRC_OVL_SHARED        ? 0
                   : variant == NRC_OVL_SHARED_DEEP ? 1
                   : variant == NRC_OVL_SHARED_PIPE ? 2
                                                    : -1;
  if (form < 0) return NRC_ERR_ARG;
  NRC_HIP(nrc::launch_split_rows((const float2*)phi, w, ns, vol, s, parts));
  NRC_HIP(nrc::launch_overlap_lds(w, (const float2*)psi, (double2*)M, ns, nt, vol, ncfg,
                                  (float*)(w + split), s, terms, form, cstride));
  return NRC_OK;
}

// times every variant `reps` times on `s` (HIP events, after one untimed call each), keeps
// the fastest in the context's table and returns it; ms[i] the best time of kOvlVariants[i]
static int overlap_tune(nrc_context* c, const void* phi, const void* psi, double* M, int ns,
                        int nt, long vol, int ncfg, void* work, hipStream_t s, int terms,
                        int reps, int* best_out, double* ms_out) {
  hipEvent_t a, b;
  NRC_HIP(hipEventCreate(&a));
  NRC_HIP(hipEventCreate(&b));
  int best = NRC_OVL_SHARED, rc = NRC_OK;
  float best_ms = -1.f;
  for (int i = 0; i < kNumOvlVariants && rc == NRC_OK; ++i) {
    const int v = kOvlVariants[i];
    float vbest = -1.f;
    for (int r = 0; r <= reps && rc == NRC_OK; ++r) {
      hipEventRecord(a, s);
      rc = overlap_bf16x3(phi, psi, M, ns, nt, vol, ncfg, work, s, terms, v);
      hipEventRecord(b, s);
      if (rc != NRC_OK) break;
      if (hipEventSynchronize(b) != hipSuccess) rc = NRC_ERR_HIP;
      float ms = 0.f;
      hipEventElapsedTime(&ms, a, b);
      if (r > 0 && (vbest < 0.f || ms < vbest)) vbest = ms;  // r = 0 is the untimed call
    }
    if (ms_out) ms_out[i] = vbest;
    if (rc == NRC_OK && (best_ms < 0.f || vbest < best_ms)) {
      best_ms = vbest;
      best = v;
    }
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (rc != NRC_OK) return rc;
  c->ovl_choice[{ns, nt, vol, terms}] = best;
  if (best_out) *best_out = best;
  return NRC_OK;
}

extern "C" {

int nrc_overlap(nrc_context* ctx, const void* phi, const void* psi, double* M, int ns, int nt,
                long vol, int ncfg, int mode) {
  if (!ctx || !phi || !psi || !M || ns < 1 || nt < 1 || vol < 1 || ncfg < 0) return NRC_ERR_ARG;
  if (mode < NRC_OVERLAP_F64 || mode > NRC_OVERLAP_BF16X6) return NRC_ERR_ARG;
  if (ncfg == 0) return NRC_OK;  // nothing to launch
  const long nM = 2L * ncfg * ns * nt;
  if (mode == NRC_OVERLAP_F64 || mode == NRC_OVERLAP_F32) {
    const int prec = mode == NRC_OVERLAP_F64 ? NRC_PREC_F64 : NRC_PREC_F32;
    NRC_HIP(nrc::launch_overlap_dense(prec, phi, psi, (double2*)M, ns, nt, vol, ncfg,
                                      ctx->stream));
    NRC_CHECK_FINITE(ctx->stream, M, nM, false, "overlap matrix");
    return NRC_OK;
  }
  if (mode != NRC_OVERLAP_BF16X3 && mode != NRC_OVERLAP_BF16X6) return NRC_ERR_ARG;
  if (ns > 64 || nt > 64 || (vol & 3)) return NRC_ERR_UNSUPPORTED;
  if (ncfg == 0) return NRC_OK;
  const int terms = mode == NRC_OVERLAP_BF16X6 ? 6 : 3;
  const int v = overlap_variant(ctx, ns, nt, vol, terms);
  void* w = ctx_workspace(ctx, overlap_bf16x3_workspace(ns, nt, vol, ncfg, v));
  if (!w) return NRC_ERR_NOMEM;
  NRC_TRY(overlap_bf16x3(phi, psi, M, ns, nt, vol, ncfg, w, ctx->stream, terms, v));
  NRC_CHECK_FINITE(ctx->stream, M, nM, false, "overlap matrix");
  return NRC_OK;
}

int nrc_overlap_set_variant(nrc_context* ctx, int ns, int nt, long vol, int mode, int variant) {
  if (!ctx || ns < 1 || nt < 1 || vol < 1) return NRC_ERR_ARG;
  if (mode != NRC_OVERLAP_BF16X3 && mode != NRC_OVERLAP_BF16X6) return NRC_ERR_ARG;
  const int terms = mode == NRC_OVERLAP_BF16X6 ? 6 : 3;
  if (variant == NRC_OVL_AUTO) {
    ctx->ovl_choice.erase({ns, nt, vol, terms});
    return NRC_OK;
  }
  if (variant < NRC_OVL_COLS || variant > NRC_OVL_SHARED_PIPE) return NRC_ERR_ARG;
  ctx->ovl_choice[{ns, nt, vol, terms}] = variant;
  return NRC_OK;
}

int nrc_overlap_variant(nrc_context* ctx, int ns, int nt, long vol, int mode, int* variant) {
  if (!ctx || !variant) return NRC_ERR_ARG;
  if (mode != NRC_OVERLAP_BF16X3 && mode != NRC_OVERLAP_BF16X6) return NRC_ERR_ARG;
  *variant = overlap_variant(ctx, ns, nt, vol, mode == NRC_OVERLAP_BF16X6 ? 6 : 3);
  return NRC_OK;
}

int nrc_overlap_tune(nrc_context* ctx, const void* phi, const void* psi, double* M, int ns,
                     int nt, long vol, int ncfg, int mode, int reps, int* best, double* ms) {
  if (!ctx || !phi || !psi || !M || ns < 1 || nt < 1 || vol < 1 || ncfg < 1 || reps < 1)
    return NRC_ERR_ARG;
  if (mode != NRC_OVERLAP_BF16X3 && mode != NRC_OVERLAP_BF16X6) return NRC_ERR_ARG;
  if (ns > 64 || nt > 64 || (vol & 3)) return NRC_ERR_UNSUPPORTED;
  void* w = ctx_workspace(ctx, overlap_bf16x3_workspace(ns, nt, vol, ncfg));
  if (!w) return NRC_ERR_NOMEM;
  return overlap_tune(ctx, phi, psi, M, ns, nt, vol, ncfg, w, ctx->stream,
                      mode == NRC_OVERLAP_BF16X6 ? 6 : 3, reps, best, ms);
}

// nrc_overlap without allocating: the split sinks and the partial tiles of the split-BF16
// modes live in the device's fixed scratch array (NRC_SCRATCH_BYTES), found at nrc_init.
// Every kernel writes the scratch before reading it, so nothing is cleared; calls that
// overlap in time on one device share the array and race.
int nrc_overlap_noalloc(nrc_context* ctx, const void* phi, const void* psi, double* M, int ns,
                        int nt, long vol, int ncfg, int mode) {
  if (!ctx || !phi || !psi || !M || ns < 1 || nt < 1 || vol < 1 || ncfg < 0) return NRC_ERR_ARG;
  if (mode < NRC_OVERLAP_F64 || mode > NRC_OVERLAP_BF16X6) return NRC_ERR_ARG;
  if (ncfg == 0) return NRC_OK;  // nothing to launch
  if (mode == NRC_OVERLAP_F64 || mode == NRC_OVERLAP_F32) {  // these use no scratch
    const int prec = mode == NRC_OVERLAP_F64 ? NRC_PREC_F64 : NRC_PREC_F32;
    NRC_HIP(nrc::launch_overlap_dense(prec, phi, psi, (double2*)M, ns, nt, vol, ncfg,
                                      ctx->stream));
    return NRC_OK;
  }
  if (mode != NRC_OVERLAP_BF16X3 && mode != NRC_OVERLAP_BF16X6) return NRC_ERR_ARG;
  if (ns > 64 || nt > 64 || (vol & 3)) return NRC_ERR_UNSUPPORTED;
  if (ncfg == 0) return NRC_OK;
  const int terms = mode == NRC_OVERLAP_BF16X6 ? 6 : 3;
  const int v = overlap_variant(ctx, ns, nt, vol, terms);
  if (overlap_bf16x3_workspace(ns, nt, vol, ncfg, v) > NRC_SCRATCH_BYTES)
    return NRC_ERR_TOO_LARGE;
  void* w = ctx->device < NRC_MAX_DEVICES ? g_scratch_dev[ctx->device] : nullptr;
  if (!w) return NRC_ERR_HIP;
  return overlap_bf16x3(phi, psi, M, ns, nt, vol, ncfg, w, ctx->stream, terms, v);
}

int nrc_pair_overlap(nrc_context* ctx, const void* psi1, const void* psi2, const double* w,
                     const int* opp, double* M, int n1, int n2, long vol, int ncfg, int mode) {
  if (!ctx || !psi1 || !psi2 || !w || !opp || !M || n1 < 1 || n2 < 1 || vol < 1 || ncfg < 0)
    return NRC_ERR_ARG;
  if (ncfg == 0) return NRC_OK;
  if (vol > INT32_MAX) return NRC_ERR_UNSUPPORTED;
  NRC_CHECK_RANGE(ctx->stream, opp, vol, 0, (int)vol, "pair_opp sites");
  const long nM = 2L * ncfg * n1 * n2;
  if (mode == NRC_OVERLAP_F64 || mode == NRC_OVERLAP_F32) {
    const int prec = mode == NRC_OVERLAP_F64 ? NRC_PREC_F64 : NRC_PREC_F32;
    NRC_HIP(nrc::launch_pair_dense(prec, psi1, psi2, w, opp, (double2*)M, n1, n2, vol, ncfg,
                                   ctx->stream));
    NRC_CHECK_FINITE(ctx->stream, M, nM, false, "pair matrix");
    return NRC_OK;
  }
  if (mode != NRC_OVERLAP_BF16X3) return NRC_ERR_ARG;
  if (n1 > 64 || n2 > 64 || (vol & 3)) return NRC_ERR_UNSUPPORTED;
  if (((uintptr_t)w | (uintptr_t)opp) & 15) return NRC_ERR_ARG;  // vector loads of w and opp
  const size_t bytes = nrc::pair_bf16x3_partial_floats(n1, n2, vol, ncfg) * sizeof(float);
  void* work = ctx_workspace(ctx, bytes);
  if (!work) return NRC_ERR_NOMEM;
  NRC_HIP(nrc::launch_pair_bf16x3((const float2*)psi1, (const float2*)psi2, w, opp, (double2*)M,
                                  n1, n2, vol, ncfg, (float*)work, ctx->stream));
  NRC_CHECK_FINITE(ctx->stream, M, nM, false, "pair matrix");
  return NRC_OK;
}

int nrc_lu_det(nrc_context* ctx, const double* M, int n, int ldm, int ncfg, int minors,
               double* logabs, double* phase, double* argsum, int* status) {
  if (!ctx || !M || !logabs || !phase || !argsum || n < 1 || ldm < n || ncfg < 0)
    return NRC_ERR_ARG;
  if (n > 64) return NRC_ERR_UNSUPPORTED;
  if (ncfg == 0) return NRC_OK;
  NRC_HIP(nrc::launch_lu_det((const double2*)M, n, ldm, ncfg, minors, logabs, (double2*)phase,
                             argsum, status, nullptr, nullptr, nullptr, ctx->stream));
  const long nout = (long)ncfg * (minors ? n : 1);
  NRC_CHECK_FINITE(ctx->stream, logabs, nout, true, "log|det|");
  NRC_CHECK_FINITE(ctx->stream, phase, 2 * nout, false, "phase");
  return NRC_OK;
}

int nrc_ludcmp(nrc_context* ctx, double* a, int n, int ncfg, int* indx, double* d, int* status) {
  if (!ctx || !a || !indx || !d || n < 1 || ncfg < 0) return NRC_ERR_ARG;
  if (n > 64) return NRC_ERR_UNSUPPORTED;
  if (ncfg == 0) return NRC_OK;
  NRC_HIP(nrc::launch_lu_det((const double2*)a, n, n, ncfg, 0, nullptr, nullptr, nullptr, status,
                             (double2*)a, indx, d, ctx->stream));
  return NRC_OK;
}

int nrc_ref_abs(nrc_context* ctx, const double* z, double* out, long n) {
  if (!ctx || !z || !out || n < 0) return NRC_ERR_ARG;
  if (n == 0) return NRC_OK;
  NRC_HIP(nrc::launch_ref_abs((const double2*)z, out, n, ctx->stream));
  return NRC_OK;
}

int nrc_z2_bits(nrc_context* ctx, uint64_t seed, uint32_t config0, int nconf, uint32_t t,
                uint32_t stream, int vol, uint32_t* words) {
  if (!ctx || !words || nconf < 0 || vol < 1) return NRC_ERR_ARG;
  if (nconf == 0) return NRC_OK;
  NRC_HIP(nrc::launch_z2_bits(seed, config0, nconf, t, stream, vol, words, ctx->stream));
  return NRC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ ensemble engine
struct nrc_ensemble {
  nrc_context* ctx = nullptr;
  hipStream_t stream = nullptr;  // the context's stream, or a batch stream while batched
  nrc_ensemble_desc d{};
  long V = 0;
  int nw = 0, nm = 1;
  std::vector<int> moms;
  void* psi = nullptr;       // [C][N][V] complex
  void* tab_x = nullptr;     // k-space split: K(1/2) = 1/xi_half; MUL_TABLE format
  void* tab_k = nullptr;     // k-space split: K = 1/xi_half^2 (two folded halves)
  double* sink_k = nullptr;  // k-space split: K(1/2)(k_s) per plane-wave sink
  double* lsrc = nullptr;    // FP32: log source_amplitude(c) per configuration
  void* sinks_k = nullptr;   // k-space split, dense sinks: K(1/2) phi_s
  void* sinks_mx = nullptr;  // FP32 k-space split, dense sinks: X-[K(1/2) phi_s] (x-chain)
  std::vector<double> khalf; // k-space split: K(1/2) per site (host)
  bool folded = false;       // the stored orbitals are xi_half psi (k-space split, t >= 1)
  bool pending_x = false;    // stored in (x, k_y, k_z) with the x-DFT(+1) pending (x-chain)
  bool has_state = false;    // psi holds orbitals in the current layout (a run has started)
  void* tab_yz = nullptr;    // x-space split: kappa/V (YZ pass); MUL_TABLE format
  void* smear = nullptr;     // k-space split + smear: f(k), MUL_TABLE format
  void* sinks = nullptr;     // dense sinks [N][V] complex (ensemble precision), or null
  void* ovl_work = nullptr;  // workspace of the split-BF16x3 overlaps of dense FP32 sinks
  bool pair = false;         // pair-sink measurement (SlaterDet3)
  double* pw = nullptr;      // pair: w(p) for the unfolded state, then w(p) K(1/2)(p) K(1/2)(opp p)
  int* popp = nullptr;       // pair: opposite-parity sites
  double2* Mp = nullptr;     // pair: [C][N][N] pair matrices
  float* pair_work = nullptr;  // pair, FP32: split-BF16x3 partial sums
  float* pws = nullptr;      // pair, FP32 partners: sg(w) sqrt|w| of both pw tables
  // dense or pair sinks on the FP32 x-chain: the folded k-space states after the slices of a
  // launch but its last, [SLICE_FUSE_MAX - 1][C][nvo][V] FP32 complex, measured after it
  void* snap = nullptr;
  size_t snap_cap = 0;       // bytes
  bool pws_neg = false;      // some pair weight is negative
  void* work = nullptr;      // reference + smear: [C][V] complex scratch
  void* rfield = nullptr;    // reference + smear: [C][V] real field
  uint32_t* words = nullptr; // [C][nw] (Philox) or [T][C][nw] (caller fields)
  size_t words_cap = 0;
  double2* M = nullptr;      // [C][N][N]
  double2* Mx = nullptr;     // [SLICE_FUSE_MAX][C][N][N]: sinks gathered inside fused launches
  double* o_la = nullptr;    // [T][C][nm]
  double2* o_ph = nullptr;
  double* o_as = nullptr;
  int* o_st = nullptr;
  size_t out_cap = 0;        // in [T][C][nm] entries
  int* g_ptr = nullptr;
  int* g_pos = nullptr;
  int* g_sink = nullptr;
  int* site_of_orb = nullptr;
  int* moms_dev = nullptr;
  // conjugate partners (nrc_ensemble_set_symmetry): with real fields and tables even under
  // p -> -p, the orbital of momentum -k is the conjugate mirror of the one of momentum k, so
  // one volume per pair is evolved; nvo volumes per configuration (N without partners)
  bool sym_ok = false;       // the kinetic (and smear) tables are even under p -> -p
  int nvo = 0;
  int* sym = nullptr;        // one device allocation holding the lists below
  const int* vo_site = nullptr;   // [nvo] source site of each volume's orbital
  const int* vo_moms = nullptr;   // [3 nvo] its momentum
  const int* vo_col = nullptr;    // [nvo] its column of M
  const int* vo_pcol = nullptr;   // [nvo] the column of its conjugate partner, or -1
  const int* g2_ptr = nullptr;    // gather lists of the mirrored sink sites -k_s
  const int* g2_pos = nullptr;
  const int* g2_sink = nullptr;
  const int* col_vol = nullptr;   // [N] volume of each column, conjugated when col_conj
  const int* col_conj = nullptr;
  const int* msite = nullptr;     // [N] site of -k_s
  std::vector<int> h_col_vol, h_col_conj;
  int mir_v0 = 0, mir_v1 = 0;  // dense sinks: the stored volumes projected on the mirrored sinks
  bool pair_sym_ok = false;  // pair sinks: opp is the mirror and w even (partners allowed)
  void* trap_tab = nullptr;  // trap (nrc_ensemble_set_trap): (1 +- A) q / V, MUL_TABLE format
  double* trap_sep = nullptr;  // a separable trap q = qx(x) qy(y) qz(z): [qx, qy, qz] (FP64)
  double* trap_q = nullptr;  // the trap q per site (FP64), for the smeared field
  double2* pgh = nullptr;    // pair sinks with partners: G then H, [C][nvo][nvo] each
  int path = NRC_PATH_AUTO;
  bool user_path = false;        // nrc_ensemble_set_path chose the path (no tuned lookup)
  int fuse = nrc::SLICE_FUSE_MAX;  // slices per launch of the persistent slice kernels
  int tile = 0;                  // x-pencils per thread of the FP64 x-chain kernel (0: default)
  std::vector<std::array<double, 3>> tune_log;  // the last tune: (kind, value, ms per slice)
  int acc = NRC_ACC_X64;         // FP32 ensembles: nrc_ensemble_set_accuracy
  uint64_t* par_dev = nullptr;   // the run's (seed, config0), read by the field kernels
  uint64_t* par_host = nullptr;  // pinned staging of par_dev
  double* tab_x64 = nullptr;     // FP32 k-space split: K(1/2) and K as FP64 tables (x-chain)
  double* tab_k64 = nullptr;
  int last_T = 0;                // slices of the last run (results are [last_T][C][nm])
  // tilted Philox fields (nrc_ensemble_set_tilt): P(+1) = tilt_thr / 2^32 per site, 0 = plain
  uint32_t tilt_thr = 0;
  int* fsum = nullptr;           // [T][C] field sums of the last tilted run
  size_t fsum_cap = 0;
  bool last_tilted = false;      // the last run drew tilted fields (fsum is valid)
  // fields with memory (nrc_ensemble_set_memory): integer parameters of launch_z2_memory
  bool mem_on = false;
  int mem_a = 0, mem_b = 0, mem_r = 0;
  int16_t* mem = nullptr;        // [C][nw * 32] per-site memory (two per site for the network)
  size_t mem_cap = 0;
  // the memory network (nrc_ensemble_set_memnet): 35 FP32 parameters; mem_on is set with it
  bool mn_on = false;
  float* mn_prm = nullptr;
  double* lrinc = nullptr;       // [T][C] log(p / q) of each slice of the last memory run
  size_t lrinc_cap = 0;
  bool last_memory = false;      // the last run drew fields with memory (lrinc is valid)
  bool profile = false;
  double prof_ms[5] = {0, 0, 0, 0, 0};
  long prof_n[5] = {0, 0, 0, 0, 0};
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_kind;
};

// A batch has captured (or will capture) this ensemble's kernels with its buffers and path
// baked in: while it belongs to a batch, nothing may reallocate its buffers, change its path,
// layout or tables, or free it (nrc_batch_destroy hands it back).
static bool ens_in_batch(const nrc_ensemble* e) { return e->stream != e->ctx->stream; }

static void ens_free(nrc_ensemble* e) {
  void* bufs[] = {e->psi, e->tab_x, e->tab_k, e->sink_k, e->lsrc, e->pw, e->popp, e->Mp,
                  e->sinks_mx,
                  e->pair_work, e->pws, e->snap, e->sinks_k, e->tab_yz, e->smear,
                  e->sinks, e->ovl_work, e->work, e->Mx,
                  e->rfield, e->words, e->M,
                  e->o_la, e->o_ph, e->o_as, e->o_st, e->g_ptr, e->g_pos, e->g_sink,
                  e->site_of_orb, e->moms_dev, e->sym, e->pgh, e->trap_tab, e->trap_q,
                  e->tab_x64, e->tab_k64, e->trap_sep, e->mem, e->lrinc, e->mn_prm};
  for (void* b : bufs)
    if (b) hipFree(b);
  if (e->par_dev) hipFree(e->par_dev);
  if (e->fsum) hipFree(e->fsum);
  if (e->par_host) hipHostFree(e->par_host);
  for (hipEvent_t x : e->ev) hipEventDestroy(x);
  delete e;
}

template <typename T>
static int upload_table(nrc_ensemble* e, void** dst, const std::vector<double>& h) {
  std::vector<T> t(h.begin(), h.end());
  if (hipMalloc(dst, t.size() * sizeof(T)) != hipSuccess) return NRC_ERR_NOMEM;
  NRC_HIP(hipMemcpy(*dst, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice));
  return NRC_OK;
}

static int ens_upload(nrc_ensemble* e, void** dst, const std::vector<double>& h) {
  return e->d.prec == NRC_PREC_F64 ? upload_table<double>(e, dst, h)
                                   : upload_table<float>(e, dst, h);
}

// a MUL_TABLE multiplier table: FP64 as is; FP32 as interleaved (nearest float, float nearest
// to the remainder) pairs
static int ens_upload_pair(nrc_ensemble* e, void** dst, const std::vector<double>& h) {
  if (e->d.prec == NRC_PREC_F64) return ens_upload(e, dst, h);
  std::vector<double> r(2 * h.size());
  for (size_t i = 0; i < h.size(); ++i) {
    r[2 * i] = (double)(float)h[i];
    r[2 * i + 1] = h[i] - r[2 * i];
  }
  return upload_table<float>(e, dst, r);
}

template <typename X>
static int dev_alloc(X** p, size_t n) {
  if (hipMalloc((void**)p, (n ? n : 1) * sizeof(X)) != hipSuccess) return NRC_ERR_NOMEM;
  return NRC_OK;
}

// s''(x, k1, k2) = sum_k0 s'(k0, k1, k2) e^{-2 pi i k0 x / L}, s' = K(1/2) s, for N complex
// sink vectors (FP64 arithmetic, twiddles by exact reduction of k0 x mod L), rounded to FP32
static std::vector<float> sinks_for_pending_x(const double* sinks, const std::vector<double>& kh,
                                              int N, int L) {
  const long P = (long)L * L, V = P * L;
  std::vector<double> c(L), sn(L);
  for (int m = 0; m < L; ++m) {
    c[m] = cos(2.0 * M_PI * m / L);
    sn[m] = -sin(2.0 * M_PI * m / L);
  }
  std::vector<float> out((size_t)2 * N * V);
  for (int n = 0; n < N; ++n)
    for (long yz = 0; yz < P; ++yz)
      for (int x = 0; x < L; ++x) {
        double re = 0.0, im = 0.0;
        for (int k0 = 0; k0 < L; ++k0) {
          const long i = (long)n * V + k0 * P + yz;
          const double a = sinks[2 * i] * kh[k0 * P + yz], b = sinks[2 * i + 1] * kh[k0 * P + yz];
          const int m = (int)(((long)k0 * x) % L);
          re += a * c[m] - b * sn[m];
          im += a * sn[m] + b * c[m];
        }
        const size_t o = (size_t)n * V + (size_t)x * P + yz;
        out[2 * o] = (float)re;
        out[2 * o + 1] = (float)im;
      }
  return out;
}

// site of -p for the site of p: (n0, n1, n2) -> ((L - n0) % L, ...), z fastest
static long mirror_site(int L, long i) {
  const long n0 = i / ((long)L * L), n1 = (i / L) % L, n2 = i % L;
  return (((L - n0) % L) * L + (L - n1) % L) * L + (L - n2) % L;
}

// the N dense sinks followed by their conjugate mirrors J phi_s, [2N][V] complex: J phi(p) =
// conj phi(-p) in k-space (conj phi(x) in x-space).  With conjugate partners the column of the
// partner J psi_r of a stored volume r is <phi_s | J psi_r> = conj <J phi_s | psi_r>.
static std::vector<double> sinks_with_mirrors(const double* sinks, int N, int L, bool kspace) {
  const long V = (long)L * L * L;
  std::vector<double> out((size_t)4 * N * V);
  std::copy(sinks, sinks + (size_t)2 * N * V, out.begin());
  for (int n = 0; n < N; ++n)
    for (long i = 0; i < V; ++i) {
      const long j = kspace ? mirror_site(L, i) : i;
      const size_t o = ((size_t)(N + n) * V + i) * 2, src = ((size_t)n * V + j) * 2;
      out[o] = sinks[src];
      out[o + 1] = -sinks[src + 1];
    }
  return out;
}

extern "C" int nrc_ensemble_create(nrc_context* ctx, const nrc_ensemble_desc* d,
                                   nrc_ensemble** out) {
  if (!ctx || !d || !out || !d->moms || !d->kin) return NRC_ERR_ARG;
  if (d->n_orb < 1 || d->n_configs < 1 || !prec_ok(d->prec)) return NRC_ERR_ARG;
  if (d->split != NRC_SPLIT_KSPACE && d->split != NRC_SPLIT_XSPACE) return NRC_ERR_ARG;
  if (!nrc::pass_supported(d->L) || d->n_orb > 64) return NRC_ERR_UNSUPPORTED;
  if (d->smear && d->split != NRC_SPLIT_KSPACE) return NRC_ERR_ARG;
  if (d->pair_w && (d->split != NRC_SPLIT_KSPACE || !d->pair_opp)) return NRC_ERR_ARG;
  const int L = d->L, N = d->n_orb;
  for (int i = 0; i < 3 * N; ++i)
    if (d->moms[i] < 0 || d->moms[i] >= L) return NRC_ERR_ARG;
  hipSetDevice(ctx->device);
  nrc_ensemble* e = new nrc_ensemble();
  e->ctx = ctx;
  e->stream = ctx->stream;
  e->d = *d;
  e->V = (long)L * L * L;
  e->nw = (int)((e->V + 31) / 32);
  e->pair = d->pair_w != nullptr;
  e->nm = e->pair ? 2 * N : (d->minors ? N : 1);
  e->moms.assign(d->moms, d->moms + 3 * N);
  e->d.moms = e->moms.data();
  e->d.kin = nullptr;
  e->d.smear = nullptr;
  e->d.sinks = nullptr;
  e->d.pair_w = nullptr;
  e->d.pair_opp = nullptr;
  const long V = e->V;
  const long C = d->n_configs;
  int rc = NRC_OK;
#define ENS_TRY(x)                 \
  do {                             \
    rc = (x);                      \
    if (rc != NRC_OK) {            \
      ens_free(e);                 \
      return rc;                   \
    }                              \
  } while (0)
  ENS_TRY(hipMalloc(&e->psi, (size_t)C * N * V * elem_bytes(d->prec)) == hipSuccess
              ? NRC_OK : NRC_ERR_NOMEM);
  ENS_TRY(dev_alloc(&e->par_dev, 2));
  ENS_TRY(hipHostMalloc((void**)&e->par_host, 2 * sizeof(uint64_t), hipHostMallocDefault) ==
                  hipSuccess ? NRC_OK : NRC_ERR_NOMEM);
  e->nvo = N;
  // tables even under p -> -p, to 1e-12 relative (sin^2(pi n / L) and sin^2(pi (L - n) / L)
  // differ in the last bit): the conjugate mirror then holds to rounding
  e->sym_ok = true;
  double smax = 0.0;
  for (long i = 0; d->smear && i < V; ++i) smax = std::max(smax, std::fabs(d->smear[i]));
  for (long i = 0; i < V && e->sym_ok; ++i) {
    const long m = mirror_site(L, i);
    const double a = d->kin[i], b = d->kin[m];
    if (!(std::fabs(a - b) <= 1e-12 * std::max(std::fabs(a), std::fabs(b)))) e->sym_ok = false;
    if (d->smear && !(std::fabs(d->smear[i] - d->smear[m]) <= 1e-12 * smax)) e->sym_ok = false;
  }
  std::vector<double> h(V);
  if (d->split == NRC_SPLIT_KSPACE) {
    e->khalf.resize(V);
    for (long i = 0; i < V; ++i) e->khalf[i] = 1.0 / d->kin[i];
    ENS_TRY(ens_upload_pair(e, &e->tab_x, e->khalf));
    for (long i = 0; i < V; ++i) h[i] = e->khalf[i] * e->khalf[i];
    ENS_TRY(ens_upload_pair(e, &e->tab_k, h));
    if (d->prec == NRC_PREC_F32) {  // the x-chain's FP64 factors
      ENS_TRY(upload_table<double>(e, (void**)&e->tab_k64, h));
      ENS_TRY(upload_table<double>(e, (void**)&e->tab_x64, e->khalf));
    }
    if (d->smear) {
      h.assign(d->smear, d->smear + V);
      ENS_TRY(ens_upload_pair(e, &e->smear, h));
      ENS_TRY(hipMalloc(&e->work, (size_t)C * V * elem_bytes(d->prec)) == hipSuccess
                  ? NRC_OK : NRC_ERR_NOMEM);
      ENS_TRY(hipMalloc(&e->rfield, (size_t)C * V * elem_bytes(d->prec) / 2) == hipSuccess
                  ? NRC_OK : NRC_ERR_NOMEM);
    }
  } else {
    for (long i = 0; i < V; ++i) h[i] = d->kin[i] / (double)V;
    ENS_TRY(ens_upload_pair(e, &e->tab_yz, h));
  }
  ENS_TRY(dev_alloc(&e->M, (size_t)C * N * N));
  ENS_TRY(dev_alloc(&e->Mx, (size_t)nrc::SLICE_FUSE_MAX * C * N * N));
  // dense sinks and their conjugate mirrors (2N sink vectors; the mirrors serve conjugate
  // partners), in the ensemble precision
  const std::vector<double> sk2 =
      d->sinks ? sinks_with_mirrors(d->sinks, N, L, d->split == NRC_SPLIT_KSPACE)
               : std::vector<double>();
  if (d->sinks) {
    if (d->prec == NRC_PREC_F64) {
      ENS_TRY(hipMalloc(&e->sinks, sk2.size() * 8) == hipSuccess ? NRC_OK : NRC_ERR_NOMEM);
      ENS_TRY(hipMemcpy(e->sinks, sk2.data(), sk2.size() * 8, hipMemcpyHostToDevice) ==
                      hipSuccess ? NRC_OK : NRC_ERR_HIP);
    } else {
      if ((V & 3) != 0) {
        ens_free(e);
        return NRC_ERR_UNSUPPORTED;
      }
      std::vector<float> f(sk2.size());
      for (size_t i = 0; i < f.size(); ++i) f[i] = (float)sk2[i];
      ENS_TRY(hipMalloc(&e->sinks, f.size() * 4) == hipSuccess ? NRC_OK : NRC_ERR_NOMEM);
      ENS_TRY(hipMemcpy(e->sinks, f.data(), f.size() * 4, hipMemcpyHostToDevice) == hipSuccess
                  ? NRC_OK : NRC_ERR_HIP);
      ENS_TRY(hipMalloc(&e->ovl_work, overlap_bf16x3_workspace(N, N, V, (int)C)) == hipSuccess
                  ? NRC_OK : NRC_ERR_NOMEM);
    }
  }
  // sink gather lists: k-space split in the X pass (pencil (k1,k2), position k0);
  // x-space split in the YZ pass column stage (pencil (k0,k2), position k1)
  const int npen = L * L;
  std::vector<int> cnt(npen + 1, 0), pos(N), snk(N), site(N);
  std::vector<std::pair<int, int>> ent(N);
  for (int s = 0; s < N; ++s) {
    const int k0 = d->moms[3 * s], k1 = d->moms[3 * s + 1], k2 = d->moms[3 * s + 2];
    site[s] = (k0 * L + k1) * L + k2;
    const int pen = d->split == NRC_SPLIT_KSPACE ? k1 * L + k2 : k0 * L + k2;
    const int p = d->split == NRC_SPLIT_KSPACE ? k0 : k1;
    ent[s] = {pen, s};
    pos[s] = p;
    cnt[pen + 1]++;
  }
  for (int p = 0; p < npen; ++p) cnt[p + 1] += cnt[p];
  std::vector<int> fill(cnt.begin(), cnt.end() - 1), gpos(N), gsink(N);
  for (int s = 0; s < N; ++s) {
    const int slot = fill[ent[s].first]++;
    gpos[slot] = pos[s];
    gsink[slot] = s;
  }
  ENS_TRY(dev_alloc(&e->g_ptr, npen + 1));
  ENS_TRY(dev_alloc(&e->g_pos, N));
  ENS_TRY(dev_alloc(&e->g_sink, N));
  ENS_TRY(dev_alloc(&e->site_of_orb, N));
  ENS_TRY(dev_alloc(&e->moms_dev, 3 * N));
  hipMemcpy(e->g_ptr, cnt.data(), (npen + 1) * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(e->g_pos, gpos.data(), N * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(e->g_sink, gsink.data(), N * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(e->site_of_orb, site.data(), N * sizeof(int), hipMemcpyHostToDevice);
  if (d->prec == NRC_PREC_F32) {  // the amplitudes of the FP32 sources, divided out by the LU
    std::vector<double> ls(C);
    for (long c = 0; c < C; ++c) ls[c] = log(nrc::source_amplitude(c));
    ENS_TRY(dev_alloc(&e->lsrc, C));
    ENS_TRY(hipMemcpy(e->lsrc, ls.data(), C * sizeof(double), hipMemcpyHostToDevice) ==
                    hipSuccess ? NRC_OK : NRC_ERR_HIP);
  }
  if (d->split == NRC_SPLIT_KSPACE) {  // K(1/2) at the sink momenta, and K(1/2)-scaled sinks
    std::vector<double> sk(N);
    for (int s2 = 0; s2 < N; ++s2) sk[s2] = e->khalf[site[s2]];
    ENS_TRY(dev_alloc(&e->sink_k, N));
    ENS_TRY(hipMemcpy(e->sink_k, sk.data(), N * sizeof(double), hipMemcpyHostToDevice) ==
                    hipSuccess ? NRC_OK : NRC_ERR_HIP);
    if (d->sinks) {
      const size_t n = (size_t)2 * N * V;
      if (d->prec == NRC_PREC_F64) {
        std::vector<double> f(2 * n);
        for (size_t i = 0; i < n; ++i) {
          const double k = e->khalf[i % V];
          f[2 * i] = sk2[2 * i] * k;
          f[2 * i + 1] = sk2[2 * i + 1] * k;
        }
        ENS_TRY(hipMalloc(&e->sinks_k, f.size() * 8) == hipSuccess ? NRC_OK : NRC_ERR_NOMEM);
        ENS_TRY(hipMemcpy(e->sinks_k, f.data(), f.size() * 8, hipMemcpyHostToDevice) ==
                        hipSuccess ? NRC_OK : NRC_ERR_HIP);
      } else {
        std::vector<float> f(2 * n);
        for (size_t i = 0; i < n; ++i) {
          const double k = e->khalf[i % V];
          f[2 * i] = (float)(sk2[2 * i] * k);
          f[2 * i + 1] = (float)(sk2[2 * i + 1] * k);
        }
        ENS_TRY(hipMalloc(&e->sinks_k, f.size() * 4) == hipSuccess ? NRC_OK : NRC_ERR_NOMEM);
        ENS_TRY(hipMemcpy(e->sinks_k, f.data(), f.size() * 4, hipMemcpyHostToDevice) ==
                        hipSuccess ? NRC_OK : NRC_ERR_HIP);
        // the same sinks for the x-chain's stored state m, whose folded k-space state is
        // X+ m (the x-DFT of sign +1 along the first index): <s|K(1/2) X+ m> =
        // <X- (K(1/2) s)|m>, since the adjoint of X+ is X-; computed here in FP64
        const std::vector<float> mx = sinks_for_pending_x(sk2.data(), e->khalf, 2 * N, L);
        ENS_TRY(hipMalloc(&e->sinks_mx, mx.size() * 4) == hipSuccess ? NRC_OK : NRC_ERR_NOMEM);
        ENS_TRY(hipMemcpy(e->sinks_mx, mx.data(), mx.size() * 4, hipMemcpyHostToDevice) ==
                        hipSuccess ? NRC_OK : NRC_ERR_HIP);
      }
    }
  }
  if (e->pair) {  // the pair wave function, unfolded (t = 0) and with K(1/2) on both momenta
    for (long i = 0; i < V; ++i)
      if (d->pair_opp[i] < 0 || d->pair_opp[i] >= V) ENS_TRY(NRC_ERR_ARG);
    // partners need opp(p) = -p and w(p) = w(-p): psi_{-t}(opp p) is then conj psi_t(p)
    double wmax = 0.0;
    for (long i = 0; i < V; ++i) wmax = std::max(wmax, std::fabs(d->pair_w[i]));
    e->pair_sym_ok = true;
    for (long i = 0; i < V && e->pair_sym_ok; ++i)
      if (d->pair_opp[i] != mirror_site(L, i) ||
          !(std::fabs(d->pair_w[i] - d->pair_w[d->pair_opp[i]]) <= 1e-12 * wmax))
        e->pair_sym_ok = false;
    std::vector<double> w(2 * V);
    for (long i = 0; i < V; ++i) {
      w[i] = d->pair_w[i];
      w[V + i] = e->khalf.empty() ? d->pair_w[i]
                                  : d->pair_w[i] * e->khalf[i] * e->khalf[d->pair_opp[i]];
    }
    ENS_TRY(dev_alloc(&e->pw, 2 * V));
    ENS_TRY(dev_alloc(&e->popp, V));
    ENS_TRY(dev_alloc(&e->Mp, (size_t)C * N * N));
    ENS_TRY(hipMemcpy(e->pw, w.data(), 2 * V * sizeof(double), hipMemcpyHostToDevice) ==
                    hipSuccess ? NRC_OK : NRC_ERR_HIP);
    ENS_TRY(hipMemcpy(e->popp, d->pair_opp, V * sizeof(int), hipMemcpyHostToDevice) ==
                    hipSuccess ? NRC_OK : NRC_ERR_HIP);
    if (d->prec == NRC_PREC_F32) {
      if (V % 4) ENS_TRY(NRC_ERR_UNSUPPORTED);
      // the N x N products, or the two products of conjugate partners in one pass
      ENS_TRY(dev_alloc(&e->pair_work, std::max(nrc::pair_bf16x3_partial_floats(N, N, V, (int)C),
                                                nrc::pair_gh_bf16x3_partial_floats(
                                                    N, V, (int)C, d->L))));
      if (e->pair_sym_ok) {  // the mirror form's scaled weights, sg(w) sqrt|w|
        std::vector<float> ws(2 * V);
        for (long i = 0; i < 2 * V; ++i) {
          ws[i] = (float)std::copysign(std::sqrt(std::fabs(w[i])), w[i]);
          if (w[i] < 0.0) e->pws_neg = true;
        }
        ENS_TRY(dev_alloc(&e->pws, 2 * V));
        ENS_TRY(hipMemcpy(e->pws, ws.data(), 2 * V * sizeof(float), hipMemcpyHostToDevice) ==
                        hipSuccess ? NRC_OK : NRC_ERR_HIP);
      }
    }
  }
  if (hipMemcpy(e->moms_dev, e->moms.data(), 3 * N * sizeof(int), hipMemcpyHostToDevice) !=
      hipSuccess) {
    ens_free(e);
    return NRC_ERR_HIP;
  }
#undef ENS_TRY
  *out = e;
  return NRC_OK;
}

extern "C" int nrc_ensemble_destroy(nrc_ensemble* e) {
  if (!e) return NRC_ERR_ARG;
  if (ens_in_batch(e)) return NRC_ERR_BATCHED;
  hipSetDevice(e->ctx->device);
  hipStreamSynchronize(e->stream);
  ens_free(e);
  return NRC_OK;
}

extern "C" int nrc_ensemble_set_path(nrc_ensemble* e, int path) {
  if (!e || path < NRC_PATH_AUTO || path > NRC_PATH_SINGLE_PASS) return NRC_ERR_ARG;
  if (ens_in_batch(e)) return NRC_ERR_BATCHED;
  if (path == NRC_PATH_SINGLE_PASS && !nrc::slice1_supported(e->d.prec, e->d.L))
    return NRC_ERR_UNSUPPORTED;
  e->path = path;
  e->user_path = path != NRC_PATH_AUTO;
  e->fuse = nrc::SLICE_FUSE_MAX;
  e->tile = 0;
  return NRC_OK;
}

extern "C" int nrc_ensemble_set_accuracy(nrc_ensemble* e, int mode) {
  if (!e || (mode != NRC_ACC_FP32 && mode != NRC_ACC_X64 && mode != NRC_ACC_XCHAIN))
    return NRC_ERR_ARG;
  if (ens_in_batch(e)) return NRC_ERR_BATCHED;
  e->acc = mode;
  return NRC_OK;
}

extern "C" int nrc_ensemble_profile(nrc_ensemble* e, int enable) {
  if (!e) return NRC_ERR_ARG;
  e->profile = enable != 0;
  for (int k = 0; k < 5; ++k) { e->prof_ms[k] = 0; e->prof_n[k] = 0; }
  return NRC_OK;
}

extern "C" int nrc_ensemble_kernel_time(nrc_ensemble* e, int kind, double* ms, long* n) {
  if (!e || kind < 0 || kind > 4) return NRC_ERR_ARG;
  if (ms) *ms = e->prof_ms[kind];
  if (n) *n = e->prof_n[kind];
  return NRC_OK;
}

// events are recorded on the ensemble's stream around each launch when profiling
struct EvScope {
  nrc_ensemble* e;
  int kind;
  size_t idx;
  EvScope(nrc_ensemble* e_, int k) : e(e_), kind(k), idx(0) {
    if (!e->profile) return;
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    idx = e->ev.size();
    e->ev.push_back(a);
    e->ev.push_back(b);
    e->ev_kind.push_back(kind);
    hipEventRecord(a, e->stream);
  }
  ~EvScope() {
    if (e->profile) hipEventRecord(e->ev[idx + 1], e->stream);
  }
};

static int ens_collect_profile(nrc_ensemble* e) {
  if (!e->profile) return NRC_OK;
  NRC_HIP(hipStreamSynchronize(e->stream));
  for (size_t i = 0; i < e->ev_kind.size(); ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e->ev[2 * i], e->ev[2 * i + 1]) == hipSuccess) {
      e->prof_ms[e->ev_kind[i]] += ms;
      e->prof_n[e->ev_kind[i]] += 1;
    }
  }
  for (hipEvent_t x : e->ev) hipEventDestroy(x);
  e->ev.clear();
  e->ev_kind.clear();
  return NRC_OK;
}

static nrc::MulSpec mul_table(const void* t) {
  nrc::MulSpec m;
  m.kind = nrc::MUL_TABLE;
  m.table = t;
  return m;
}

// the reference field (1 + F-[f F+ phi])/V, computed as F+[f F-[phi]] (equal for the even,
// real f of periodic lattices), real part kept: src/hamiltonian/noarch/hamiltonian_initialize.C
template <typename T>
__global__ void k_bits_to_complex(nrc::C2<T>* out, const uint32_t* words, int nw, long V,
                                  long C) {
  const long n = C * V;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long c = i / V, s = i - c * V;
    const uint32_t w = words[c * nw + (s >> 5)];
    out[i] = {((w >> (s & 31)) & 1u) ? T(1) : T(-1), T(0)};
  }
}

// the real field (re + 1)/V, times the trap q(site) when one is set (q null: 1)
template <typename T>
__global__ void k_complex_to_field(const nrc::C2<T>* in, T* out, long n, double invV,
                                   const double* q, long V) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const double f = ((double)in[i].x + 1.0) * invV;
    out[i] = (T)(q ? f * q[i % V] : f);
  }
}

static int ens_smear_field(nrc_ensemble* e, const uint32_t* words) {
  const long n = (long)e->d.n_configs * e->V;
  long blocks = std::min<long>(65536, (n + 255) / 256);
  hipStream_t s = e->stream;
  if (e->d.prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_bits_to_complex<double>, dim3(blocks), dim3(256), 0, s,
                       (nrc::C2<double>*)e->work, words, e->nw, e->V, (long)e->d.n_configs);
  else
    hipLaunchKernelGGL(k_bits_to_complex<float>, dim3(blocks), dim3(256), 0, s,
                       (nrc::C2<float>*)e->work, words, e->nw, e->V, (long)e->d.n_configs);
  NRC_HIP(hipGetLastError());
  nrc::PassArgs a;
  a.psi = e->work;
  a.nvol = e->d.n_configs;
  a.L = e->d.L;
  a.fft_b = 1;
  NRC_HIP(nrc::launch_xpass(e->d.prec, a, s));
  a.fft_a = 1;
  a.fft_b = 1;
  a.m1 = mul_table(e->smear);
  NRC_HIP(nrc::launch_yzpass(e->d.prec, a, s));
  nrc::PassArgs b;
  b.psi = e->work;
  b.nvol = e->d.n_configs;
  b.L = e->d.L;
  b.fft_a = 1;
  NRC_HIP(nrc::launch_xpass(e->d.prec, b, s));
  if (e->d.prec == NRC_PREC_F64)
    hipLaunchKernelGGL(k_complex_to_field<double>, dim3(blocks), dim3(256), 0, s,
                       (const nrc::C2<double>*)e->work, (double*)e->rfield, n,
                       1.0 / (double)e->V, (const double*)e->trap_q, e->V);
  else
    hipLaunchKernelGGL(k_complex_to_field<float>, dim3(blocks), dim3(256), 0, s,
                       (const nrc::C2<float>*)e->work, (float*)e->rfield, n,
                       1.0 / (double)e->V, (const double*)e->trap_q, e->V);
  NRC_HIP(hipGetLastError());
  return NRC_OK;
}

static int ens_reserve_outputs(nrc_ensemble* e, int T) {
  const size_t need = (size_t)T * e->d.n_configs * e->nm;
  if (need <= e->out_cap) return NRC_OK;
  hipFree(e->o_la); hipFree(e->o_ph); hipFree(e->o_as); hipFree(e->o_st);
  e->o_la = nullptr; e->o_ph = nullptr; e->o_as = nullptr; e->o_st = nullptr;
  e->out_cap = 0;
  NRC_TRY(dev_alloc(&e->o_la, need));
  NRC_TRY(dev_alloc(&e->o_ph, need));
  NRC_TRY(dev_alloc(&e->o_as, need));
  NRC_TRY(dev_alloc(&e->o_st, need));
  e->out_cap = need;
  return NRC_OK;
}

static nrc::GatherSpec ens_gather_spec(const nrc_ensemble* e) {
  nrc::GatherSpec g;
  g.ptr = e->g_ptr;
  g.pos = e->g_pos;
  g.sink = e->g_sink;
  g.M = e->M;
  g.nsink = e->d.n_orb;
  g.scale = e->d.split == NRC_SPLIT_KSPACE ? 1.0 : 1.0 / sqrt((double)e->V);
  if (e->sym) {
    g.col = e->vo_col;
    g.pcol = e->vo_pcol;
    g.ptr2 = e->g2_ptr;
    g.pos2 = e->g2_pos;
    g.sink2 = e->g2_sink;
    g.ncol = e->d.n_orb;
  }
  return g;
}

// Propagator::Set(one_bodies[n]->Get(), n): plane-wave sources in every configuration
static int ens_set_sources(nrc_ensemble* e) {
  const long C = e->d.n_configs;
  e->folded = false;
  e->pending_x = false;
  e->has_state = true;
  const int nvo = e->nvo;
  if (e->d.split == NRC_SPLIT_KSPACE)
    NRC_HIP(nrc::launch_set_unit_sources(e->d.prec, e->psi, e->sym ? e->vo_site : e->site_of_orb,
                                         C * nvo, nvo, e->V, e->stream));
  else
    NRC_HIP(nrc::launch_set_plane_waves(e->d.prec, e->psi, e->sym ? e->vo_moms : e->moms_dev,
                                        C * nvo, nvo, e->d.L, e->stream));
  return NRC_OK;
}

// ---- one slice of the single-pass path, in two halves (the loop of ens_enqueue, and the
// step entry points that let a caller choose each slice's field after seeing the measurement)

// LU log-dets of the measurement M (the one-body overlaps [C][N][N]) into the results row
// `row`: the full matrix or its leading minors; in pair mode SlaterDet3's 2N determinants
// (production/slater_det3_trapped/main-cum.C:217-238), m = 2n + 1 from the pair matrices
// e->Mp, m = 2n the same minor with column n taken from M.  FP32 sources carry the
// amplitude r_c: r_c^2 per pair column, r_c per one-body column.
static int ens_lu(nrc_ensemble* e, const double2* M, size_t row) {
  const int N = e->d.n_orb;
  const int C = (int)e->d.n_configs;
  nrc::LuOpts o;
  o.lsrc = e->lsrc;
  if (!e->pair) {
    NRC_HIP(nrc::launch_lu_det(M, N, N, C, e->d.minors, e->o_la + row, e->o_ph + row,
                               e->o_as + row, e->o_st + row, nullptr, nullptr, nullptr, e->stream,
                               o));
    return NRC_OK;
  }
  o.lsrc_cols = 2;
  o.onm = 2 * N;
  o.ostep = 2;
  o.oofs = 1;
  NRC_HIP(nrc::launch_lu_det(e->Mp, N, N, C, 1, e->o_la + row, e->o_ph + row, e->o_as + row,
                             e->o_st + row, nullptr, nullptr, nullptr, e->stream, o));
  o.oofs = 0;
  o.repl = M;
  o.ldr = N;
  NRC_HIP(nrc::launch_lu_det(e->Mp, N, N, C, 1, e->o_la + row, e->o_ph + row, e->o_as + row,
                             e->o_st + row, nullptr, nullptr, nullptr, e->stream, o));
  return NRC_OK;
}

// measurement at t into the results row t: overlaps (gathered plane-wave sinks or dense
// sinks) and the LU log-det (production/slater_det/main.C:206-215); in pair mode also the
// pair matrices of the stored orbitals (TwoBody::Run for every (s, t),
// src/slater_det/slater_det3.C:115-122)
// psi_in: measure that folded k-space state instead of the stored one (the states a launch
// of several slices wrote for the slices inside it)
static int ens_measure(nrc_ensemble* e, int t, const void* psi_in = nullptr) {
  hipStream_t s = e->stream;
  const int N = e->d.n_orb, prec = e->d.prec;
  const long C = e->d.n_configs, V = e->V;
  const size_t per_t = (size_t)C * e->nm;
  void* const psi = psi_in ? const_cast<void*>(psi_in) : e->psi;
  const bool folded = psi_in ? true : e->folded;
  const bool pending_x = psi_in ? false : e->pending_x;
  {
    EvScope sc(e, 4);
    // folded k-space state xi_half psi: sinks carry the missing K(1/2); a state with its
    // x-DFT(+1) pending: the sinks transformed by X- as well
    const void* sk = pending_x ? e->sinks_mx : folded ? e->sinks_k : e->sinks;
    if (e->sinks && e->sym) {
      // conjugate partners: the sinks and their mirrors against the stored volumes, then the
      // N x N matrix (a partner's column is the conjugate of its mirrored-sink overlaps)
      // (the mirrored sinks only against the volumes that have a partner: a contiguous range
      // for shells filled in the reference's momentum order, so k = 0 drops out)
      const int nv = e->nvo, v0 = e->mir_v0, nm = e->mir_v1 - e->mir_v0;
      const size_t esz = prec == NRC_PREC_F64 ? 16 : 8;
      double2* Mp = e->Mx;
      double2* Mm = e->Mx + (size_t)C * N * nv;
      const void* skm = (const char*)sk + (size_t)N * V * esz;
      const void* pm = (const char*)psi + (size_t)v0 * V * esz;
      if (prec == NRC_PREC_F64) {
        NRC_HIP(nrc::launch_overlap_dense(prec, sk, psi, Mp, N, nv, V, (int)C, s));
        NRC_HIP(nrc::launch_overlap_dense(prec, skm, pm, Mm, N, nm, V, (int)C, s, nv * V));
      } else {
        const int terms = e->acc != NRC_ACC_FP32 ? 6 : 3;
        NRC_TRY(overlap_bf16x3(sk, psi, (double*)Mp, N, nv, V, (int)C, e->ovl_work, s, terms,
                               overlap_variant(e->ctx, N, nv, V, terms)));
        NRC_TRY(overlap_bf16x3(skm, pm, (double*)Mm, N, nm, V, (int)C, e->ovl_work, s, terms,
                               overlap_variant(e->ctx, N, nm, V, terms), nv * V));
      }
      NRC_HIP(nrc::launch_sink_expand(Mp, Mm, e->col_vol, e->col_conj, e->M, N, nv, nm, v0, C,
                                      s));
    } else if (e->sinks && prec == NRC_PREC_F64)
      NRC_HIP(nrc::launch_overlap_dense(prec, sk, psi, e->M, N, N, V, (int)C, s));
    else if (e->sinks) {
      const int terms = e->acc != NRC_ACC_FP32 ? 6 : 3;
      NRC_TRY(overlap_bf16x3(sk, psi, (double*)e->M, N, N, V, (int)C, e->ovl_work, s, terms,
                             overlap_variant(e->ctx, N, N, V, terms)));
    }
    else {
      nrc::ConjCols cp;
      if (e->sym) {
        cp.vol = e->col_vol;
        cp.conj = e->col_conj;
        cp.msite = e->msite;
        cp.nvo = e->nvo;
      }
      NRC_HIP(nrc::launch_gather_sites(prec, psi, e->site_of_orb,
                                       folded ? e->sink_k : nullptr, e->M, C, N, V, s, cp));
    }
    if (e->pair && e->sym) {  // G (mirrored sites) and H (conjugated) of the stored volumes
      const double* w = e->pw + (folded ? V : 0);
      const int nv = e->nvo;
      double2* G = e->pgh;
      double2* H = e->pgh + (size_t)C * nv * nv;
      if (prec == NRC_PREC_F64) {
        NRC_HIP(nrc::launch_pair_dense(prec, psi, psi, w, e->popp, G, nv, nv, V, (int)C, s));
        NRC_HIP(nrc::launch_pair_dense(prec, psi, psi, w, nullptr, H, nv, nv, V, (int)C, s));
      } else if (nv <= 48) {  // both products in one pass
        // the mirror form reads each site and its mirror together (pair_opp(p) = -p and an
        // even w are conditions of partners; L a multiple of 4 and at least 8)
        const int Lm = (e->d.L % 4 == 0 && e->d.L >= 8) ? e->d.L : 0;
        NRC_HIP(nrc::launch_pair_gh_bf16x3((const float2*)psi, w, e->popp, G, H, nv, V,
                                           (int)C, e->pair_work, s, e->pws ? Lm : 0,
                                           e->pws ? e->pws + (folded ? V : 0) : nullptr,
                                           e->pws_neg));
      } else {
        NRC_HIP(nrc::launch_pair_bf16x3((const float2*)psi, (const float2*)psi, w,
                                        e->popp, G, nv, nv, V, (int)C, e->pair_work, s));
        NRC_HIP(nrc::launch_pair_bf16x3((const float2*)psi, (const float2*)psi, w,
                                        nullptr, H, nv, nv, V, (int)C, e->pair_work, s));
      }
      NRC_HIP(nrc::launch_pair_expand(G, H, e->col_vol, e->col_conj, e->Mp, N, nv, C, s));
    } else if (e->pair) {
      const double* w = e->pw + (folded ? V : 0);
      if (prec == NRC_PREC_F64)
        NRC_HIP(nrc::launch_pair_dense(prec, psi, psi, w, e->popp, e->Mp, N, N, V,
                                       (int)C, s));
      else
        NRC_HIP(nrc::launch_pair_bf16x3((const float2*)psi, (const float2*)psi, w,
                                        e->popp, e->Mp, N, N, V, (int)C, e->pair_work, s));
    }
  }
  EvScope sc(e, 2);
  return ens_lu(e, e->M, t * per_t);
}

// evolution of slice t with the fields words[c][nw] (device), single-pass kernel; for the
// x-space split with gathered sinks the measurement of t happens inside the slice
static int ens_evolve(nrc_ensemble* e, int t, const uint32_t* words) {
  hipStream_t s = e->stream;
  const int N = e->d.n_orb, L = e->d.L, prec = e->d.prec;
  const long C = e->d.n_configs, V = e->V;
  const bool ref = e->d.split == NRC_SPLIT_KSPACE;
  const size_t per_t = (size_t)C * e->nm;
  if (ref && e->smear) NRC_TRY(ens_smear_field(e, words));
  nrc::MulSpec fieldspec;
  fieldspec.kind = nrc::MUL_FIELD;
  fieldspec.nw = e->nw;
  fieldspec.bits = words;
  nrc::SliceArgs a;
  a.psi = e->psi;
  a.nvol = C * e->nvo;
  a.norb = e->nvo;
  a.L = L;
  a.grid = e->ctx->n_cu;  // persistent kernel: workgroups per CU x CUs
  if (ref) {
    a.pre = mul_table(e->folded ? e->tab_k : e->tab_x);  // post: folded into the next pre
    if (e->smear) {
      a.mid.kind = nrc::MUL_RFIELD;
      a.mid.table = e->rfield;
    } else {
      a.mid = fieldspec;
      a.mid.a0 = 1.0 / (double)V;
      a.mid.a1 = e->d.A / (double)V;
      a.mid.table = e->trap_tab;  // with a trap, the factors per site (null without)
    }
  } else {
    a.mid = mul_table(e->tab_yz);
    if (!e->sinks) {
      a.gather_mid = 1;
      a.g = ens_gather_spec(e);
    }
    a.post = fieldspec;
    a.post.a0 = 1.0;
    a.post.a1 = e->d.A;
  }
  {
    EvScope sc(e, 0);
    NRC_HIP(nrc::launch_slice1(prec, a, s));
  }
  e->folded = ref;
  if (!ref && !e->sinks) {  // measured inside the slice
    EvScope sc(e, 2);
    NRC_TRY(ens_lu(e, e->M, t * per_t));
  }
  return NRC_OK;
}

// slices t .. t + S - 1 in one launch of the persistent kernel (k-space split, plain field):
// words[k * stride ..] are slice t + k's fields; the sinks gathered after slice t + k
// (k < S - 1) are factored into result row t + k + 1
static int ens_evolve_fused(nrc_ensemble* e, int t, int S, const uint32_t* words, size_t stride) {
  hipStream_t s = e->stream;
  const int N = e->d.n_orb;
  const long C = e->d.n_configs, V = e->V;
  const size_t per_t = (size_t)C * e->nm;
  const long mstride = C * N * N;
  nrc::SliceArgs a;
  a.psi = e->psi;
  a.nvol = C * e->nvo;
  a.norb = e->nvo;
  a.L = e->d.L;
  a.grid = e->ctx->n_cu;
  a.pre = mul_table(e->folded ? e->tab_k : e->tab_x);
  a.mid.kind = nrc::MUL_FIELD;
  a.mid.nw = e->nw;
  a.mid.bits = words;
  a.mid.a0 = 1.0 / (double)V;
  a.mid.a1 = e->d.A / (double)V;
  a.mid.table = e->trap_tab;  // null without a trap
  a.nslices = S;
  a.pre_next = e->tab_k;
  a.bits_stride = (long)stride;
  a.m_stride = mstride;
  a.g = ens_gather_spec(e);
  a.g.M = e->Mx;
  a.g.sink_scale = e->sink_k;  // the state after each slice is the folded one
  {
    EvScope sc(e, 0);
    NRC_HIP(nrc::launch_slice1(e->d.prec, a, s));
  }
  e->folded = true;
  if (S > 1) {  // the S - 1 measurements gathered inside the launch: one LU launch for all
    EvScope sc(e, 2);
    nrc::LuOpts o;
    o.lsrc = e->lsrc;
    o.lsrc_period = C;
    const size_t row = (size_t)(t + 1) * per_t;  // result rows t + 1 .. t + S - 1, contiguous
    NRC_HIP(nrc::launch_lu_det(e->Mx, N, N, (int)(C * (S - 1)), e->d.minors, e->o_la + row,
                               e->o_ph + row, e->o_as + row, e->o_st + row, nullptr, nullptr,
                               nullptr, s, o));
  }
  return NRC_OK;
}

// slices t .. t + S - 1 in one launch of the x-chain kernel (slice_ext.hip), with their
// measurements gathered inside it into Mx and one LU launch for them: the launch starts from
// the sources (t = 0) or the state the previous launch left with its x-DFT(+1) pending, and
// the run's last launch stores the folded k-space state
static bool ens_xchain(const nrc_ensemble* e, bool single) {
  return single && e->d.prec == NRC_PREC_F32 && e->acc != NRC_ACC_FP32 &&
         e->d.split == NRC_SPLIT_KSPACE && !e->smear && (!e->trap_tab || e->trap_sep) &&
         nrc::slicex_supported(e->d.L);
}

// gathered plane-wave sinks: measured inside the launch (S slices), the state left with its
// x-DFT(+1) pending between launches; dense sinks: measured beforehand by ens_measure on that
// state, one slice per launch; pair sinks: measured beforehand from the folded k-space state,
// which every launch (one slice) starts from and stores
static int ens_evolve_x(nrc_ensemble* e, int t, int S, const uint32_t* words, size_t stride,
                        bool last) {
  const bool gather = e->sinks == nullptr && !e->pair;
  const bool kspace = e->pair;  // every launch starts and ends in k-space
  const int N = e->d.n_orb;
  const long C = e->d.n_configs, V = e->V;
  const size_t per_t = (size_t)C * e->nm;
  nrc::SliceXArgs a;
  a.psi = e->psi;
  a.nvol = C * e->nvo;
  a.norb = e->nvo;
  a.tile = e->tile;
  a.zd = e->acc == NRC_ACC_X64;
  a.L = e->d.L;
  a.grid = e->ctx->n_cu;
  a.nslices = S;
  a.start_k = t == 0 || kspace;
  a.half0 = t == 0;
  a.store_k = last || kspace;
  a.trap_sep = e->trap_sep;
  a.tab_x = e->tab_x64;
  a.tab_k = e->tab_k64;
  a.bits = words;
  a.nw = e->nw;
  a.bits_stride = (long)stride;
  a.a0 = 1.0 / (double)V;
  a.a1 = e->d.A / (double)V;
  a.gather = gather;
  a.g = ens_gather_spec(e);
  a.g.M = e->Mx;
  a.g.sink_scale = e->sink_k;  // not on the sources (t = 0): the kernel drops it there
  a.m_stride = C * N * N;
  if (!gather && S > 1) {  // the states inside the launch, for ens_measure
    if (!e->snap || e->snap_cap < (size_t)(S - 1) * C * e->nvo * V * 8) return NRC_ERR_ARG;
    a.snap = e->snap;
    a.snap_stride = C * e->nvo * V;
  }
  {
    EvScope sc(e, 0);
    NRC_HIP(nrc::launch_slicex(a, e->stream));
  }
  e->folded = last || kspace;
  e->pending_x = !(last || kspace);
  if (!gather) return NRC_OK;
  EvScope sc(e, 2);
  nrc::LuOpts o;
  o.lsrc = e->lsrc;
  o.lsrc_period = C;
  const size_t row = (size_t)t * per_t;  // result rows t .. t + S - 1, contiguous
  NRC_HIP(nrc::launch_lu_det(e->Mx, N, N, (int)(C * S), e->d.minors, e->o_la + row,
                             e->o_ph + row, e->o_as + row, e->o_st + row, nullptr, nullptr,
                             nullptr, e->stream, o));
  return NRC_OK;
}

// allocations and host-to-device copies of a run of T slices (caller fields: words_host)
static int ens_prepare(nrc_ensemble* e, const uint32_t* words_host, int T) {
  if (T < 1) return NRC_ERR_ARG;
  hipSetDevice(e->ctx->device);
  // dense or pair sinks on the FP32 x-chain: room for the states inside a launch of several
  // slices (without it, or where it does not fit, such runs take one slice per launch)
  if ((e->sinks || e->pair) && e->d.prec == NRC_PREC_F32 && e->acc != NRC_ACC_FP32 &&
      e->d.split == NRC_SPLIT_KSPACE && !e->smear && nrc::slicex_supported(e->d.L) && T > 1 &&
      e->fuse > 1) {
    const size_t need = (size_t)(nrc::SLICE_FUSE_MAX - 1) * e->d.n_configs * e->nvo * e->V * 8;
    if (need > e->snap_cap) {
      NRC_HIP(hipStreamSynchronize(e->stream));
      hipFree(e->snap);
      e->snap = nullptr;
      e->snap_cap = 0;
      if (hipMalloc(&e->snap, need) == hipSuccess) e->snap_cap = need;
      else (void)hipGetLastError();
    }
  }
  const size_t slice_words = (size_t)e->d.n_configs * e->nw;
  const size_t need_words = words_host ? slice_words * T : slice_words * nrc::SLICE_FUSE_MAX;
  if (need_words > e->words_cap) {
    NRC_HIP(hipStreamSynchronize(e->stream));
    hipFree(e->words);
    e->words = nullptr;
    e->words_cap = 0;
    NRC_TRY(dev_alloc(&e->words, need_words));
    e->words_cap = need_words;
  }
  if (words_host)
    NRC_HIP(hipMemcpyAsync(e->words, words_host, need_words * sizeof(uint32_t),
                           hipMemcpyHostToDevice, e->stream));
  if (e->mem_on && !words_host) {
    const size_t need_mem = (size_t)e->d.n_configs * e->nw * 32 * (e->mn_on ? 2 : 1);
    if (need_mem > e->mem_cap) {
      NRC_HIP(hipStreamSynchronize(e->stream));
      hipFree(e->mem);
      e->mem = nullptr;
      e->mem_cap = 0;
      NRC_TRY(dev_alloc(&e->mem, need_mem));
      e->mem_cap = need_mem;
    }
    const size_t need_lr = (size_t)T * e->d.n_configs;
    if (need_lr > e->lrinc_cap) {
      NRC_HIP(hipStreamSynchronize(e->stream));
      hipFree(e->lrinc);
      e->lrinc = nullptr;
      e->lrinc_cap = 0;
      NRC_TRY(dev_alloc(&e->lrinc, need_lr));
      e->lrinc_cap = need_lr;
    }
  }
  const size_t need_sums = e->tilt_thr && !words_host ? (size_t)T * e->d.n_configs : 0;
  if (need_sums > e->fsum_cap) {
    NRC_HIP(hipStreamSynchronize(e->stream));
    hipFree(e->fsum);
    e->fsum = nullptr;
    e->fsum_cap = 0;
    NRC_TRY(dev_alloc(&e->fsum, need_sums));
    e->fsum_cap = need_sums;
  }
  return ens_reserve_outputs(e, T);
}

// the run's seed and first configuration into the device parameter block the field kernels
// read, on `s` (the ensemble's stream, or the stream a batch launches its graph on); the
// previous run has completed (runs synchronise), so the pinned staging words are free
static int ens_set_params(nrc_ensemble* e, uint64_t seed, uint32_t config0, hipStream_t s) {
  e->par_host[0] = seed;
  e->par_host[1] = config0;
  NRC_HIP(hipMemcpyAsync(e->par_dev, e->par_host, 2 * sizeof(uint64_t), hipMemcpyHostToDevice,
                         s));
  return NRC_OK;
}

// the kernels of a run of T slices on the ensemble's stream; nothing here allocates, copies
// to or from the host or synchronises, so a batch can capture it into a graph
static int ens_enqueue(nrc_ensemble* e, uint64_t seed, uint32_t config0, bool host_words, int T) {
  hipStream_t s = e->stream;
  const int N = e->d.n_orb, L = e->d.L, prec = e->d.prec;
  const long C = e->d.n_configs, V = e->V;
  const bool ref = e->d.split == NRC_SPLIT_KSPACE;
  const size_t slice_words = (size_t)C * e->nw;
  const size_t per_t = (size_t)C * e->nm;

  NRC_TRY(ens_set_sources(e));
  const nrc::GatherSpec g = ens_gather_spec(e);
  // Lattice::Refresh for slice tt into dst: the plain sampler's Philox bits, tilted ones with
  // their field sums, or fields with memory with their sampling ratios
  auto refresh = [&](int tt, uint32_t* dst) -> hipError_t {
    if (e->mn_on)
      return nrc::launch_z2_memnet(seed, config0, (int)C, (uint32_t)tt, (int)V, e->mn_prm, dst,
                                   reinterpret_cast<uint32_t*>(e->mem),
                                   e->lrinc + (size_t)tt * C, s, e->par_dev);
    if (e->mem_on)
      return nrc::launch_z2_memory(seed, config0, (int)C, (uint32_t)tt, NRC_STREAM_MEMORY,
                                   (int)V, e->mem_a, e->mem_b, e->mem_r, dst, e->mem,
                                   e->lrinc + (size_t)tt * C, s, e->par_dev);
    if (e->tilt_thr)
      return nrc::launch_z2_tilt(seed, config0, (int)C, (uint32_t)tt, NRC_STREAM_TILT, (int)V,
                                 e->tilt_thr, dst, e->fsum + (size_t)tt * C, s, e->par_dev);
    return nrc::launch_z2_bits(seed, config0, (int)C, (uint32_t)tt, 0u, (int)V, dst, s,
                               e->par_dev);
  };

  auto field_of = [&](int t) -> const uint32_t* {
    return host_words ? e->words + (size_t)t * slice_words : e->words;
  };
  auto lu = [&](int t) -> int {
    EvScope sc(e, 2);
    return ens_lu(e, e->M, t * per_t);
  };
  nrc::MulSpec fieldspec;
  fieldspec.kind = nrc::MUL_FIELD;
  fieldspec.nw = e->nw;

  const bool single = e->path == NRC_PATH_SINGLE_PASS ||
                      (e->path == NRC_PATH_AUTO && nrc::slice1_supported(prec, L));
  // single pass: every slice is one kernel over the orbitals (slice_fused.hip)
  // dense sinks and pair matrices read the stored state
  if ((e->sinks || e->pair) && !single) return NRC_ERR_UNSUPPORTED;
  if (ens_xchain(e, single)) {  // FP32 state, FP64 x-chain, measurements inside the launches
    e->pending_x = false;
    const bool outside = e->sinks || e->pair;  // dense or pair sinks: measured outside
    const size_t snap_need = (size_t)(nrc::SLICE_FUSE_MAX - 1) * C * e->nvo * V * 8;
    const int fo = outside && (e->snap == nullptr || e->snap_cap < snap_need) ? 1 : e->fuse;
    for (int t = 0; t < T;) {
      const int S = std::min(fo, T - t);
      if (outside) NRC_TRY(ens_measure(e, t));  // the state the launch starts from
      if (!host_words) {
        EvScope sc(e, 3);
        for (int k = 0; k < S; ++k) NRC_HIP(refresh(t + k, e->words + (size_t)k * slice_words));
      }
      NRC_TRY(ens_evolve_x(e, t, S, host_words ? field_of(t) : e->words, slice_words,
                           t + S == T));
      for (int k = 1; outside && k < S; ++k)  // the states the launch wrote for t + 1 ..
        NRC_TRY(ens_measure(e, t + k, (const char*)e->snap + (size_t)(k - 1) * C * e->nvo * V * 8));
      t += S;
    }
    return NRC_OK;
  }
  const bool measure_first = ref || e->sinks;
  // k-space split with the plain field: up to SLICE_FUSE_MAX slices per launch of the
  // persistent kernel, the measurements between them gathered inside it
  const bool fuse = single && ref && !e->smear && !e->sinks && !e->pair &&
                    (!e->trap_tab || nrc::slice1p_trap_ok(prec, L));
  for (int t = 0; single && t < T;) {
    const int S = fuse ? std::min(e->fuse, T - t) : 1;
    if (measure_first) NRC_TRY(ens_measure(e, t));
    if (!host_words) {
      EvScope sc(e, 3);
      for (int k = 0; k < S; ++k) NRC_HIP(refresh(t + k, e->words + (size_t)k * slice_words));
    }
    if (S > 1)
      NRC_TRY(ens_evolve_fused(e, t, S, host_words ? field_of(t) : e->words, slice_words));
    else
      NRC_TRY(ens_evolve(e, t, field_of(t)));
    t += S;
  }

  // two passes per slice: an X pass that finishes slice t-1, measures t and starts slice t,
  // then a YZ pass
  for (int t = 0; !single && t <= T; ++t) {
    const bool last = t == T;
    // X pass: finish slice t-1, measure t (k-space split), start slice t
    {
      nrc::PassArgs a;
      a.psi = e->psi;
      a.nvol = C * e->nvo;
      a.norb = e->nvo;
      a.L = L;
      a.fft_a = t > 0;
      a.fft_b = !last;
      if (ref) {  // the trailing K(1/2) of slice t-1 is folded into the next factor
        if (!last) {
          a.gather = 1;
          a.g = g;
          if (t > 0) a.g.sink_scale = e->sink_k;
          a.m2 = mul_table(t > 0 ? e->tab_k : e->tab_x);
        }
      } else if (t > 0) {
        a.m1 = fieldspec;
        a.m1.bits = field_of(t - 1);
        a.m1.a0 = 1.0;
        a.m1.a1 = e->d.A;
      }
      EvScope sc(e, 0);
      NRC_HIP(nrc::launch_xpass(prec, a, s));
    }
    if (last) break;
    if (ref) NRC_TRY(lu(t));
    // Lattice::Refresh for slice t
    if (!host_words) {
      EvScope sc(e, 3);
      NRC_HIP(refresh(t, e->words));
    }
    if (ref && e->smear) NRC_TRY(ens_smear_field(e, field_of(t)));
    // YZ pass: the field (k-space split) or kinetic (x-space split) multiply between the FFTs
    {
      nrc::PassArgs a;
      a.psi = e->psi;
      a.nvol = C * e->nvo;
      a.norb = e->nvo;
      a.L = L;
      a.fft_a = 1;
      a.fft_b = 1;
      if (ref) {
        if (e->smear) {
          a.m1.kind = nrc::MUL_RFIELD;
          a.m1.table = e->rfield;
        } else {
          a.m1 = fieldspec;
          a.m1.bits = field_of(t);
          a.m1.a0 = 1.0 / (double)V;
          a.m1.a1 = e->d.A / (double)V;
          a.m1.table = e->trap_tab;  // null without a trap
        }
      } else {
        a.gather = 1;
        a.g = g;
        a.m1 = mul_table(e->tab_yz);
      }
      EvScope sc(e, 1);
      NRC_HIP(nrc::launch_yzpass(prec, a, s));
    }
    if (!ref) NRC_TRY(lu(t));
  }
  if (!single) e->folded = ref;
  return NRC_OK;
}

// the per-slice results of the last run, [t][c][m], copied to the host (any may be null)
static int ens_collect(nrc_ensemble* e, int T, double* logabs, double* phase, double* argsum) {
  hipStream_t s = e->stream;
  const size_t tot = (size_t)e->d.n_configs * e->nm * T;
  NRC_CHECK_FINITE(s, e->o_la, (long)tot, true, "ensemble log|det|");
  NRC_CHECK_FINITE(s, (const double*)e->o_ph, 2 * (long)tot, false, "ensemble phase");
  if (logabs) NRC_HIP(hipMemcpyAsync(logabs, e->o_la, tot * sizeof(double), hipMemcpyDeviceToHost, s));
  if (phase) NRC_HIP(hipMemcpyAsync(phase, e->o_ph, tot * sizeof(double2), hipMemcpyDeviceToHost, s));
  if (argsum) NRC_HIP(hipMemcpyAsync(argsum, e->o_as, tot * sizeof(double), hipMemcpyDeviceToHost, s));
  NRC_HIP(hipStreamSynchronize(s));
  return ens_collect_profile(e);
}

// the key of an ensemble's tuned choice: shape, precision, layout and measurement kind
static std::array<long, 8> ens_tune_key(const nrc_ensemble* e) {
  const long flags = (e->sinks ? 1 : 0) | (e->pair ? 2 : 0) | (e->trap_tab ? 4 : 0) |
                     (e->smear ? 8 : 0);
  return {e->d.L, e->d.n_orb, e->d.prec, e->d.split, e->nvo, e->acc, flags, e->d.n_configs};
}

// a tuned (path, slices per launch) for the ensemble's key, unless the caller chose the path
static void ens_apply_tuned(nrc_ensemble* e) {
  if (e->user_path || ens_in_batch(e)) return;
  const auto it = e->ctx->ens_choice.find(ens_tune_key(e));
  if (it == e->ctx->ens_choice.end()) return;
  e->path = it->second[0];
  e->fuse = it->second[1];
  e->tile = it->second[2];
}

static int ens_run(nrc_ensemble* e, uint64_t seed, uint32_t config0, const uint32_t* words_host,
                   int T, double* logabs, double* phase, double* argsum) {
  ens_apply_tuned(e);
  NRC_TRY(ens_prepare(e, words_host, T));
  NRC_TRY(ens_set_params(e, seed, config0, e->stream));
  NRC_TRY(ens_enqueue(e, seed, config0, words_host != nullptr, T));
  e->last_T = T;
  e->last_tilted = e->tilt_thr && !words_host && !e->mem_on;
  e->last_memory = e->mem_on && !words_host;
  return ens_collect(e, T, logabs, phase, argsum);
}

extern "C" int nrc_ensemble_run(nrc_ensemble* e, uint64_t seed, uint32_t config0, int T,
                                double* logabs, double* phase, double* argsum) {
  if (!e) return NRC_ERR_ARG;
  if (ens_in_batch(e)) return NRC_ERR_BATCHED;
  return ens_run(e, seed, config0, nullptr, T, logabs, phase, argsum);
}

extern "C" int nrc_ensemble_run_fields(nrc_ensemble* e, const uint32_t* words, int T,
                                       double* logabs, double* phase, double* argsum) {
  if (!e || !words) return NRC_ERR_ARG;
  if (ens_in_batch(e)) return NRC_ERR_BATCHED;
  return ens_run(e, 0, 0, words, T, logabs, phase, argsum);
}

// the orbitals of configuration c: the stored volumes, with the conjugate mirrors of the
// partners filled in (psi_{-k}(p) = conj psi_k(-p) in k-space, conj psi_k(x) in x-space)
template <typename T>
static void ens_expand_partners(const nrc_ensemble* e, T* o) {
  const int N = e->d.n_orb, nvo = e->nvo;
  const long V = e->V;
  const bool kspace = e->d.split == NRC_SPLIT_KSPACE;
  std::vector<T> vol((size_t)2 * nvo * V);
  std::copy(o, o + vol.size(), vol.begin());
  for (int n = 0; n < N; ++n) {
    const T* src = vol.data() + (size_t)2 * e->h_col_vol[n] * V;
    T* dst = o + (size_t)2 * n * V;
    const bool cj = e->h_col_conj[n] != 0;
    for (long i = 0; i < V; ++i) {
      const long j = cj && kspace ? mirror_site(e->d.L, i) : i;
      dst[2 * i] = src[2 * j];
      dst[2 * i + 1] = cj ? -src[2 * j + 1] : src[2 * j + 1];
    }
  }
}

extern "C" int nrc_ensemble_orbitals(nrc_ensemble* e, int c, void* out) {
  if (!e || !out || c < 0 || c >= e->d.n_configs) return NRC_ERR_ARG;
  if (!e->has_state) return NRC_ERR_ARG;  // no run since creation or the last layout change
  const size_t vbytes = (size_t)e->V * elem_bytes(e->d.prec);
  NRC_HIP(hipMemcpyAsync(out, (char*)e->psi + (size_t)c * e->nvo * vbytes, e->nvo * vbytes,
                         hipMemcpyDeviceToHost, e->stream));
  NRC_HIP(hipStreamSynchronize(e->stream));
  if (e->sym) {
    if (e->d.prec == NRC_PREC_F64) ens_expand_partners(e, (double*)out);
    else ens_expand_partners(e, (float*)out);
  }
  const size_t n = (size_t)e->d.n_orb * e->V;
  if (e->d.prec == NRC_PREC_F64) {
    if (e->folded) {  // stored xi_half psi: apply the slice's trailing K(1/2)
      double* o = (double*)out;
      for (size_t i = 0; i < n; ++i) {
        o[2 * i] *= e->khalf[i % e->V];
        o[2 * i + 1] *= e->khalf[i % e->V];
      }
    }
  } else {  // and divide out the configuration's source amplitude
    const double r = 1.0 / nrc::source_amplitude(c);
    float* o = (float*)out;
    for (size_t i = 0; i < n; ++i) {
      const double f = e->folded ? e->khalf[i % e->V] * r : r;
      o[2 * i] = (float)((double)o[2 * i] * f);
      o[2 * i + 1] = (float)((double)o[2 * i + 1] * f);
    }
  }
  return NRC_OK;
}

// Conjugate partners: with a real field and kinetic (and smear) tables even under p -> -p,
// the evolution commutes with psi(p) -> conj psi(-p), and the plane-wave source of momentum
// -k is the image of the one of momentum k; so only one orbital of each (k, -k) pair is
// evolved and the other's sink values are read conjugated at the mirrored sinks.  Orbital
// n is a representative unless an earlier orbital already claimed it as its partner.
// Trap: every slice the x-space multiplier of the field becomes (1 + A phi) q / V (or the
// smeared field times q).  For the trapped production order K, P(1/2), I, P(1/2), K
// (production/slater_det3_trapped/main-cum.C:97-106) q = (V p)^2 with p Potential's table
// (src/hamiltonian/hamiltonian.C:163-240): the three x-space multiplies commute, and each
// forward / backward pair with its 1/V is the multiply alone.
extern "C" int nrc_ensemble_set_trap(nrc_ensemble* e, const double* q) {
  if (!e) return NRC_ERR_ARG;
  if (ens_in_batch(e)) return NRC_ERR_BATCHED;
  if (q && e->d.split != NRC_SPLIT_KSPACE) return NRC_ERR_UNSUPPORTED;
  const long V = e->V;
  for (long i = 0; q && i < V; ++i)
    if (!std::isfinite(q[i])) return NRC_ERR_ARG;
  hipSetDevice(e->ctx->device);
  NRC_HIP(hipStreamSynchronize(e->stream));
  if (e->trap_tab) hipFree(e->trap_tab);
  if (e->trap_q) hipFree(e->trap_q);
  if (e->trap_sep) hipFree(e->trap_sep);
  e->trap_tab = nullptr;
  e->trap_q = nullptr;
  e->trap_sep = nullptr;
  if (!q) return NRC_OK;
  {  // separable, q(x,y,z) = qx(x) qy(y) qz(z) to 1e-14 relative (the harmonic EXP trap is):
     // qx = q(x,0,0)/q(0,0,0), qy = q(0,y,0)/q(0,0,0), qz = q(0,0,z)
    const int L = e->d.L;
    const long P = (long)L * L;
    const double q0 = q[0];
    std::vector<double> sep(3 * (size_t)L);
    for (int i = 0; i < L; ++i) {
      sep[i] = q[i * P] / q0;
      sep[L + i] = q[i * L] / q0;
      sep[2 * L + i] = q[i];
    }
    bool ok = q0 != 0.0;
    for (long i = 0; i < V && ok; ++i) {
      const double r = sep[i / P] * sep[L + (i / L) % L] * sep[2 * L + i % L];
      ok = std::fabs(r - q[i]) <= 1e-14 * std::fabs(q[i]);
    }
    if (ok) {
      NRC_TRY(dev_alloc(&e->trap_sep, sep.size()));
      NRC_HIP(hipMemcpy(e->trap_sep, sep.data(), sep.size() * sizeof(double),
                        hipMemcpyHostToDevice));
    }
  }
  std::vector<double> h(2 * V);
  const double a0 = 1.0 / (double)V, a1 = e->d.A / (double)V;
  for (long i = 0; i < V; ++i) {
    h[i] = (a0 + a1) * q[i];
    h[V + i] = (a0 - a1) * q[i];
  }
  NRC_TRY(ens_upload_pair(e, &e->trap_tab, h));
  NRC_TRY(dev_alloc(&e->trap_q, V));
  NRC_HIP(hipMemcpy(e->trap_q, q, V * sizeof(double), hipMemcpyHostToDevice));
  return NRC_OK;
}

extern "C" int nrc_ensemble_volumes(nrc_ensemble* e, int* nvo) {
  if (!e || !nvo) return NRC_ERR_ARG;
  *nvo = e->nvo;
  return NRC_OK;
}

extern "C" int nrc_ensemble_set_symmetry(nrc_ensemble* e, int conj) {
  if (!e || (conj != 0 && conj != 1)) return NRC_ERR_ARG;
  if (ens_in_batch(e)) return NRC_ERR_BATCHED;
  if (conj && (!e->sym_ok || (e->pair && !e->pair_sym_ok))) return NRC_ERR_UNSUPPORTED;
  hipSetDevice(e->ctx->device);
  NRC_HIP(hipStreamSynchronize(e->stream));
  if (e->sym) hipFree(e->sym);
  if (e->pgh) hipFree(e->pgh);
  e->sym = nullptr;
  e->pgh = nullptr;
  e->has_state = false;  // the stored volumes no longer match the layout
  e->nvo = e->d.n_orb;
  e->h_col_vol.clear();
  e->h_col_conj.clear();
  if (!conj) return NRC_OK;
  const int N = e->d.n_orb, L = e->d.L;
  const int* k = e->moms.data();
  auto site_of = [&](int a, int b, int c2) { return (a * L + b) * L + c2; };
  auto neg = [&](int a) { return (L - a) % L; };
  std::vector<int> col_vol(N, -1), col_conj(N, 0), vo_col, vo_pcol, vo_site, vo_moms;
  for (int n = 0; n < N; ++n) {
    if (col_vol[n] >= 0) continue;  // the partner of an earlier representative
    const int r = (int)vo_col.size();
    col_vol[n] = r;
    const bool self = neg(k[3 * n]) == k[3 * n] && neg(k[3 * n + 1]) == k[3 * n + 1] &&
                      neg(k[3 * n + 2]) == k[3 * n + 2];
    int pc = -1;
    for (int m = n + 1; m < N && !self && pc < 0; ++m)
      if (col_vol[m] < 0 && k[3 * m] == neg(k[3 * n]) && k[3 * m + 1] == neg(k[3 * n + 1]) &&
          k[3 * m + 2] == neg(k[3 * n + 2]))
        pc = m;
    if (pc >= 0) {
      col_vol[pc] = r;
      col_conj[pc] = 1;
    }
    vo_col.push_back(n);
    vo_pcol.push_back(pc);
    vo_site.push_back(site_of(k[3 * n], k[3 * n + 1], k[3 * n + 2]));
    for (int j = 0; j < 3; ++j) vo_moms.push_back(k[3 * n + j]);
  }
  const int nvo = (int)vo_col.size();
  // the volumes with a partner: [v0, v1) when they are contiguous, else every volume
  int v0 = nvo, v1 = 0;
  for (int r = 0; r < nvo; ++r)
    if (vo_pcol[r] >= 0) {
      v0 = std::min(v0, r);
      v1 = std::max(v1, r + 1);
    }
  bool contiguous = v1 > v0;
  for (int r = v0; r < v1 && contiguous; ++r) contiguous = vo_pcol[r] >= 0;
  if (!contiguous) {
    v0 = 0;
    v1 = nvo;
  }
  // gather lists of the mirrored sinks, in the pencil layout of the plain ones: k-space split
  // x-pencils (pencil (k1, k2), position k0), x-space split y-columns (pencil (k0, k2),
  // position k1)
  const int npen = L * L;
  std::vector<int> cnt(npen + 1, 0), pen(N), pos(N), msite(N);
  for (int s = 0; s < N; ++s) {
    const int m0 = neg(k[3 * s]), m1 = neg(k[3 * s + 1]), m2 = neg(k[3 * s + 2]);
    msite[s] = site_of(m0, m1, m2);
    pen[s] = e->d.split == NRC_SPLIT_KSPACE ? m1 * L + m2 : m0 * L + m2;
    pos[s] = e->d.split == NRC_SPLIT_KSPACE ? m0 : m1;
    cnt[pen[s] + 1]++;
  }
  for (int p = 0; p < npen; ++p) cnt[p + 1] += cnt[p];
  std::vector<int> fill(cnt.begin(), cnt.end() - 1), gpos(N), gsink(N);
  for (int s = 0; s < N; ++s) {
    const int slot = fill[pen[s]]++;
    gpos[slot] = pos[s];
    gsink[slot] = s;
  }
  std::vector<int> all;
  auto put = [&](const std::vector<int>& v) {
    const size_t o = all.size();
    all.insert(all.end(), v.begin(), v.end());
    return o;
  };
  const size_t o_site = put(vo_site), o_moms = put(vo_moms), o_col = put(vo_col),
               o_pcol = put(vo_pcol), o_ptr = put(cnt), o_pos = put(gpos), o_sink = put(gsink),
               o_cv = put(col_vol), o_cc = put(col_conj), o_ms = put(msite);
  int* d = nullptr;
  NRC_TRY(dev_alloc(&d, all.size()));
  if (hipMemcpy(d, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    hipFree(d);
    return NRC_ERR_HIP;
  }
  if (e->pair) {
    const int rc = dev_alloc(&e->pgh, (size_t)2 * e->d.n_configs * nvo * nvo);
    if (rc != NRC_OK) {
      hipFree(d);
      return rc;
    }
  }
  e->sym = d;
  e->vo_site = d + o_site;
  e->vo_moms = d + o_moms;
  e->vo_col = d + o_col;
  e->vo_pcol = d + o_pcol;
  e->g2_ptr = d + o_ptr;
  e->g2_pos = d + o_pos;
  e->g2_sink = d + o_sink;
  e->col_vol = d + o_cv;
  e->col_conj = d + o_cc;
  e->msite = d + o_ms;
  e->nvo = nvo;
  e->h_col_vol = col_vol;
  e->h_col_conj = col_conj;
  e->mir_v0 = v0;
  e->mir_v1 = v1;
  return NRC_OK;
}

// ------------------------------------------------------------------ tilted fields
extern "C" int nrc_ensemble_set_tilt(nrc_ensemble* e, double p_plus) {
  if (!e || !(p_plus > 0.0 && p_plus < 1.0)) return NRC_ERR_ARG;
  if (ens_in_batch(e)) return NRC_ERR_BATCHED;
  e->mem_on = e->mn_on = false;  // one importance sampler at a time
  if (p_plus == 0.5) {
    e->tilt_thr = 0;
    return NRC_OK;
  }
  const double thr = std::nearbyint(std::ldexp(p_plus, 32));
  e->tilt_thr = (uint32_t)std::min(std::max(thr, 1.0), 4294967295.0);
  return NRC_OK;
}

extern "C" int nrc_ensemble_tilt(nrc_ensemble* e, double* p_plus) {
  if (!e || !p_plus) return NRC_ERR_ARG;
  *p_plus = e->tilt_thr ? std::ldexp((double)e->tilt_thr, -32) : 0.5;
  return NRC_OK;
}

// ------------------------------------------------------------------ fields with memory
// (lambda, theta, rho) to the integer parameters of launch_z2_memory: a = 2^31 lambda,
// b = 2^21 theta (M = 1024 m), r = 2^15 rho, each rounded to nearest
static int mem_params(double lambda, double theta, double rho, int* a, int* b, int* r) {
  if (!(std::fabs(lambda) <= 0x1p-6) || !(std::fabs(theta) <= 0x1p-6) ||
      !(rho >= 0.0 && rho <= 0.96))
    return NRC_ERR_ARG;
  *a = (int)std::nearbyint(std::ldexp(lambda, 31));
  *b = (int)std::nearbyint(std::ldexp(theta, 21));
  *r = (int)std::nearbyint(std::ldexp(rho, 15));
  return NRC_OK;
}

extern "C" int nrc_ensemble_set_memory(nrc_ensemble* e, double lambda, double theta,
                                       double rho) {
  if (!e) return NRC_ERR_ARG;
  int a, b, r;
  NRC_TRY(mem_params(lambda, theta, rho, &a, &b, &r));
  if (ens_in_batch(e)) return NRC_ERR_BATCHED;
  e->mem_a = a;
  e->mem_b = b;
  e->mem_r = r;
  e->mem_on = a != 0 || b != 0;
  e->mn_on = false;
  if (e->mem_on) e->tilt_thr = 0;  // one importance sampler at a time
  return NRC_OK;
}

// ------------------------------------------------------------------ the memory network
static constexpr int kMnParams = 35;

static bool mn_params_ok(const float* p) {
  for (int i = 0; i < kMnParams; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}

extern "C" int nrc_ensemble_set_memnet(nrc_ensemble* e, const float* params) {
  if (!e) return NRC_ERR_ARG;
  if (params && !mn_params_ok(params)) return NRC_ERR_ARG;
  if (ens_in_batch(e)) return NRC_ERR_BATCHED;
  if (!params) {
    if (e->mn_on) e->mem_on = false;
    e->mn_on = false;
    return NRC_OK;
  }
  if (!e->mn_prm && hipMalloc(&e->mn_prm, kMnParams * sizeof(float)) != hipSuccess) {
    e->mn_prm = nullptr;
    return NRC_ERR_NOMEM;
  }
  NRC_HIP(hipStreamSynchronize(e->stream));
  NRC_HIP(hipMemcpy(e->mn_prm, params, kMnParams * sizeof(float), hipMemcpyHostToDevice));
  e->mn_on = e->mem_on = true;
  e->mem_a = e->mem_b = 0;
  e->tilt_thr = 0;  // one importance sampler at a time
  return NRC_OK;
}

// the memory network on the plain sampler's fields: grad false, out_dev[t][c] = cumulative
// log(p / q); grad true, out_dev[c][35] = the gradient of sum_t G[t][c] sum log q
static int mn_walk(nrc_context* ctx, uint64_t seed, uint32_t config0, int nconf, int T, int vol,
                   const float* params, const double* G_dev, double* out_dev, bool grad) {
  if (!ctx || !params || !out_dev || (grad && !G_dev) || nconf < 0 || T < 0 || vol < 1)
    return NRC_ERR_ARG;
  if (!mn_params_ok(params)) return NRC_ERR_ARG;
  if (((long)vol + 31) / 32 * 32 > 32768) return NRC_ERR_UNSUPPORTED;
  if (nconf == 0 || T == 0) return NRC_OK;
  hipSetDevice(ctx->device);
  float* prm = (float*)ctx_workspace(ctx, kMnParams * sizeof(float));
  if (!prm) return NRC_ERR_NOMEM;
  NRC_HIP(hipMemcpyAsync(prm, params, kMnParams * sizeof(float), hipMemcpyHostToDevice,
                         ctx->stream));
  NRC_HIP(nrc::launch_z2_memnet_walk(seed, config0, nconf, T, vol, prm, G_dev, out_dev, grad,
                                     ctx->stream));
  NRC_HIP(hipStreamSynchronize(ctx->stream));  // the parameters are read before returning
  return NRC_OK;
}

extern "C" int nrc_z2_memnet_eval(nrc_context* ctx, uint64_t seed, uint32_t config0, int nconf,
                                  int T, int vol, const float* params, double* out_dev) {
  return mn_walk(ctx, seed, config0, nconf, T, vol, params, nullptr, out_dev, false);
}

extern "C" int nrc_z2_memnet_grad(nrc_context* ctx, uint64_t seed, uint32_t config0, int nconf,
                                  int T, int vol, const float* params, const double* G_dev,
                                  double* grad_dev) {
  return mn_walk(ctx, seed, config0, nconf, T, vol, params, G_dev, grad_dev, true);
}

extern "C" int nrc_ensemble_memory(nrc_ensemble* e, double* lambda, double* theta,
                                   double* rho) {
  if (!e || !lambda || !theta || !rho) return NRC_ERR_ARG;
  *lambda = e->mem_on ? std::ldexp((double)e->mem_a, -31) : 0.0;
  *theta = e->mem_on ? std::ldexp((double)e->mem_b, -21) : 0.0;
  *rho = std::ldexp((double)e->mem_r, -15);
  return NRC_OK;
}

extern "C" int nrc_z2_memory_eval(nrc_context* ctx, uint64_t seed, uint32_t config0, int nconf,
                                  int T, int vol, double lambda, double theta, double rho,
                                  double* out_dev) {
  if (!ctx || !out_dev || nconf < 0 || T < 0 || vol < 1) return NRC_ERR_ARG;
  if (((long)vol + 31) / 32 * 32 > 32768) return NRC_ERR_UNSUPPORTED;
  int a, b, r;
  NRC_TRY(mem_params(lambda, theta, rho, &a, &b, &r));
  if (nconf == 0 || T == 0) return NRC_OK;
  hipSetDevice(ctx->device);
  NRC_HIP(nrc::launch_z2_memory_eval(seed, config0, nconf, T, vol, a, b, r, out_dev,
                                     ctx->stream));
  return NRC_OK;
}

// log r_t[c] = sum over the slices t' < t of log(p(phi_t') / q(phi_t')): p the plain
// sampler's 2^-V, q the tilted one's q^n(+1) (1 - q)^n(-1); zero for a plain run
extern "C" int nrc_ensemble_log_ratio(nrc_ensemble* e, int T, double* out) {
  if (!e || !out || e->last_T < 1 || T != e->last_T) return NRC_ERR_ARG;
  const long C = e->d.n_configs, V = e->V;
  if (e->last_memory) {
    std::vector<double> inc((size_t)T * C);
    NRC_HIP(hipStreamSynchronize(e->stream));
    NRC_HIP(hipMemcpy(inc.data(), e->lrinc, inc.size() * sizeof(double),
                      hipMemcpyDeviceToHost));
    for (long c = 0; c < C; ++c) out[c] = 0.0;
    for (int t = 1; t < T; ++t)
      for (long c = 0; c < C; ++c)
        out[(size_t)t * C + c] = out[(size_t)(t - 1) * C + c] + inc[(size_t)(t - 1) * C + c];
    return NRC_OK;
  }
  if (!e->last_tilted) {
    std::fill(out, out + (size_t)T * C, 0.0);
    return NRC_OK;
  }
  std::vector<int> sums((size_t)T * C);
  NRC_HIP(hipStreamSynchronize(e->stream));
  NRC_HIP(hipMemcpy(sums.data(), e->fsum, sums.size() * sizeof(int), hipMemcpyDeviceToHost));
  const double q = std::ldexp((double)e->tilt_thr, -32);
  const double lp = std::log(0.5 / q), lm = std::log(0.5 / (1.0 - q));
  for (long c = 0; c < C; ++c) out[c] = 0.0;
  for (int t = 1; t < T; ++t)
    for (long c = 0; c < C; ++c) {
      const long np = (V + sums[(size_t)(t - 1) * C + c]) / 2;
      out[(size_t)t * C + c] = out[(size_t)(t - 1) * C + c] + np * lp + (V - np) * lm;
    }
  return NRC_OK;
}

extern "C" int nrc_ensemble_results(nrc_ensemble* e, double* logabs, double* phase,
                                    double* argsum) {
  if (!e || e->last_T < 1) return NRC_ERR_ARG;
  return ens_collect(e, e->last_T, logabs, phase, argsum);
}

// ------------------------------------------------------------------ autotuner
// Times the ensemble's variants over T_probe slices (HIP events on the ensemble's stream,
// after one untimed slice): the evolution paths its shape supports; with the path chosen,
// the slices per launch of the persistent kernels (1, 2 or 4, where the path fuses slices);
// with dense FP32 sinks, the overlap GEMM variants on the ensemble's own buffers.  Keeps the
// fastest of each, records every timing (nrc_ensemble_tune_report) and stores the choice in
// the context under the ensemble's key (nrc_tune_save / nrc_tune_load), from where later
// runs of ensembles with the same key take it.
static int ens_time_slices(nrc_ensemble* e, int T_probe, hipEvent_t a, hipEvent_t b, float* ms) {
  NRC_TRY(ens_run(e, 0x5eedULL, 0xfffff000u, nullptr, 1, nullptr, nullptr, nullptr));
  NRC_TRY(ens_prepare(e, nullptr, T_probe));
  NRC_TRY(ens_set_params(e, 0x5eedULL, 0xfffff000u, e->stream));
  hipEventRecord(a, e->stream);
  NRC_TRY(ens_enqueue(e, 0x5eedULL, 0xfffff000u, false, T_probe));
  hipEventRecord(b, e->stream);
  if (hipEventSynchronize(b) != hipSuccess) return NRC_ERR_HIP;
  hipEventElapsedTime(ms, a, b);
  return NRC_OK;
}

extern "C" int nrc_ensemble_tune(nrc_ensemble* e, int T_probe, int* path_out,
                                 double* ms_per_slice) {
  if (!e || T_probe < 1) return NRC_ERR_ARG;
  if (ens_in_batch(e)) return NRC_ERR_BATCHED;
  std::vector<int> cand = {NRC_PATH_TWO_PASS};
  if (nrc::slice1_supported(e->d.prec, e->d.L)) cand.push_back(NRC_PATH_SINGLE_PASS);
  if (e->sinks || e->pair) cand = {NRC_PATH_SINGLE_PASS};
  hipEvent_t a, b;
  NRC_HIP(hipEventCreate(&a));
  NRC_HIP(hipEventCreate(&b));
  e->tune_log.clear();
  const bool user_path = e->user_path;
  e->user_path = true;  // no tuned lookup while timing
  int best = cand[0], best_fuse = nrc::SLICE_FUSE_MAX, rc = NRC_OK;
  float best_ms = -1.f;
  e->fuse = nrc::SLICE_FUSE_MAX;
  for (int p : cand) {
    e->path = p;
    float ms = 0.f;
    rc = ens_time_slices(e, T_probe, a, b, &ms);
    if (rc != NRC_OK) break;
    e->tune_log.push_back({0.0, (double)p, ms / T_probe});
    if (best_ms < 0.f || ms < best_ms) {
      best_ms = ms;
      best = p;
    }
  }
  e->path = best;
  // slices per launch: only where the chosen path fuses slices into one launch (dense and
  // pair sinks: on the FP32 x-chain, which writes the states inside a launch for them)
  const bool fuses = rc == NRC_OK && best == NRC_PATH_SINGLE_PASS &&
                     e->d.split == NRC_SPLIT_KSPACE && !e->smear &&
                     (!(e->sinks || e->pair) || ens_xchain(e, true));
  if (fuses) {
    const int Tf = std::max(T_probe, nrc::SLICE_FUSE_MAX);
    float base = -1.f;
    for (int f : {nrc::SLICE_FUSE_MAX, 2, 1}) {
      e->fuse = f;
      float ms = 0.f;
      rc = ens_time_slices(e, Tf, a, b, &ms);
      if (rc != NRC_OK) break;
      e->tune_log.push_back({1.0, (double)f, ms / Tf});
      if (base < 0.f || ms < base) {
        base = ms;
        best_fuse = f;
      }
    }
    if (rc == NRC_OK) best_ms = base * T_probe / Tf;
  }
  e->fuse = best_fuse;
  // x-pencils per thread of the FP64 x-chain kernel, where the shape offers more than one
  int best_tile = 0;
  e->tile = 0;
  int tiles[4];
  const int ntiles = nrc::slicex_tiles(e->d.L, tiles);
  if (rc == NRC_OK && ntiles > 1 && ens_xchain(e, best == NRC_PATH_SINGLE_PASS ||
                                                     (best == NRC_PATH_AUTO &&
                                                      nrc::slice1_supported(e->d.prec,
                                                                            e->d.L)))) {
    const int Tt = std::max(T_probe, e->fuse);
    float tb = -1.f;
    for (int i = 0; i < ntiles && rc == NRC_OK; ++i) {
      e->tile = tiles[i];
      float ms = 0.f;
      rc = ens_time_slices(e, Tt, a, b, &ms);
      if (rc != NRC_OK) break;
      e->tune_log.push_back({3.0, (double)tiles[i], ms / Tt});
      if (tb < 0.f || ms < tb) {
        tb = ms;
        best_tile = tiles[i];
      }
    }
    if (rc == NRC_OK) best_ms = tb * T_probe / Tt;
  }
  e->tile = best_tile;
  // dense FP32 sinks: the overlap GEMM variants on the ensemble's buffers
  if (rc == NRC_OK && e->sinks && e->d.prec == NRC_PREC_F32) {
    const int terms = e->acc != NRC_ACC_FP32 ? 6 : 3;
    const int N = e->d.n_orb;
    double ms[kNumOvlVariants];
    int v = 0;
    rc = overlap_tune(e->ctx, e->sinks, e->psi, (double*)e->Mx, N, e->sym ? e->nvo : N, e->V,
                      (int)e->d.n_configs, e->ovl_work, e->stream, terms, 2, &v, ms);
    for (int i = 0; rc == NRC_OK && i < kNumOvlVariants; ++i)
      e->tune_log.push_back({2.0, (double)kOvlVariants[i], ms[i]});
    const int nm = e->mir_v1 - e->mir_v0;
    if (rc == NRC_OK && e->sym && nm != e->nvo) {  // the mirrored sinks' shape as well
      rc = overlap_tune(e->ctx, e->sinks, e->psi, (double*)e->Mx, N, nm, e->V,
                        (int)e->d.n_configs, e->ovl_work, e->stream, terms, 2, &v, ms);
      for (int i = 0; rc == NRC_OK && i < kNumOvlVariants; ++i)
        e->tune_log.push_back({2.0, (double)kOvlVariants[i], ms[i]});
    }
    if (rc == NRC_OK) {  // the slice time with the chosen GEMM
      float t = 0.f;
      rc = ens_time_slices(e, T_probe, a, b, &t);
      best_ms = t;
    }
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  e->user_path = user_path;
  e->last_T = 0;
  if (rc != NRC_OK) return rc;
  e->ctx->ens_choice[ens_tune_key(e)] = {best, best_fuse, best_tile};
  if (path_out) *path_out = best;
  if (ms_per_slice) *ms_per_slice = best_ms / T_probe;
  return NRC_OK;
}

extern "C" int nrc_ensemble_tune_report(nrc_ensemble* e, int i, int* kind, int* value,
                                        double* ms) {
  if (!e || i < 0) return NRC_ERR_ARG;
  if ((size_t)i >= e->tune_log.size()) return NRC_ERR_ARG;
  if (kind) *kind = (int)e->tune_log[i][0];
  if (value) *value = (int)e->tune_log[i][1];
  if (ms) *ms = e->tune_log[i][2];
  return NRC_OK;
}

extern "C" int nrc_ensemble_launch_slices(nrc_ensemble* e, int* slices) {
  if (!e || !slices) return NRC_ERR_ARG;
  ens_apply_tuned(e);
  *slices = e->fuse;
  return NRC_OK;
}

// ------------------------------------------------------------------ tuned choices on disk
// A text file: a header line, then one line per choice,
//   overlap <ns> <nt> <vol> <terms> <variant>
//   ensemble <L> <N> <prec> <split> <volumes> <accuracy> <flags> <configs> <path> <slices>
//            <tile>
// (version 1 files, without the tile, load with the default tile)
extern "C" int nrc_tune_save(nrc_context* ctx, const char* path) {
  if (!ctx || !path) return NRC_ERR_ARG;
  FILE* f = fopen(path, "w");
  if (!f) return NRC_ERR_ARG;
  fprintf(f, "nrcps-tune 2\n");
  for (const auto& kv : ctx->ovl_choice)
    fprintf(f, "overlap %ld %ld %ld %ld %d\n", kv.first[0], kv.first[1], kv.first[2],
            kv.first[3], kv.second);
  for (const auto& kv : ctx->ens_choice) {
    fprintf(f, "ensemble");
    for (long v : kv.first) fprintf(f, " %ld", v);
    fprintf(f, " %d %d %d\n", kv.second[0], kv.second[1], kv.second[2]);
  }
  return fclose(f) == 0 ? NRC_OK : NRC_ERR_ARG;
}

extern "C" int nrc_tune_load(nrc_context* ctx, const char* path) {
  if (!ctx || !path) return NRC_ERR_ARG;
  FILE* f = fopen(path, "r");
  if (!f) return NRC_ERR_ARG;
  char tag[32];
  int version = 0;
  int rc = (fscanf(f, "%31s %d", tag, &version) == 2 && !strcmp(tag, "nrcps-tune") &&
            (version == 1 || version == 2))
               ? NRC_OK
               : NRC_ERR_ARG;
  std::map<std::array<long, 4>, int> ovl;
  std::map<std::array<long, 8>, std::array<int, 3>> ens;
  while (rc == NRC_OK && fscanf(f, "%31s", tag) == 1) {
    if (!strcmp(tag, "overlap")) {
      std::array<long, 4> k;
      int v;
      if (fscanf(f, "%ld %ld %ld %ld %d", &k[0], &k[1], &k[2], &k[3], &v) != 5 ||
          v < NRC_OVL_COLS || v > NRC_OVL_SHARED_PIPE)
        rc = NRC_ERR_ARG;
      else
        ovl[k] = v;
    } else if (!strcmp(tag, "ensemble")) {
      std::array<long, 8> k;
      std::array<int, 3> v = {0, 0, 0};
      for (long& x : k)
        if (fscanf(f, "%ld", &x) != 1) rc = NRC_ERR_ARG;
      if (rc == NRC_OK && (fscanf(f, "%d %d", &v[0], &v[1]) != 2 || v[0] < NRC_PATH_AUTO ||
                           v[0] > NRC_PATH_SINGLE_PASS || v[1] < 1 ||
                           v[1] > nrc::SLICE_FUSE_MAX))
        rc = NRC_ERR_ARG;
      if (rc == NRC_OK && version == 2 && (fscanf(f, "%d", &v[2]) != 1 || v[2] < 0 || v[2] > 2))
        rc = NRC_ERR_ARG;
      if (rc == NRC_OK) ens[k] = v;
    } else {
      rc = NRC_ERR_ARG;
    }
  }
  fclose(f);
  if (rc != NRC_OK) return rc;  // a malformed file changes nothing
  for (const auto& kv : ovl) ctx->ovl_choice[kv.first] = kv.second;
  for (const auto& kv : ens) ctx->ens_choice[kv.first] = kv.second;
  return NRC_OK;
}

// ------------------------------------------------------------------ batches of shape groups
struct nrc_batch {
  nrc_context* ctx = nullptr;
  std::vector<nrc_ensemble*> g;
  std::vector<hipStream_t> streams;  // one per group, default priority
  std::vector<hipEvent_t> done;      // one per group (fork/join)
  hipEvent_t fork = nullptr;
  hipGraphExec_t exec = nullptr;
  int exec_T = 0;
  long captures = 0;  // graphs captured and instantiated so far
  double capture_ms = 0.0;
};

static void batch_drop_graph(nrc_batch* b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  b->exec_T = 0;
}

extern "C" int nrc_batch_create(nrc_context* ctx, nrc_ensemble* const* groups, int n_groups,
                                nrc_batch** out) {
  if (!ctx || !groups || n_groups < 1 || !out) return NRC_ERR_ARG;
  for (int i = 0; i < n_groups; ++i) {
    if (!groups[i] || groups[i]->ctx != ctx) return NRC_ERR_ARG;
    if (ens_in_batch(groups[i])) return NRC_ERR_BATCHED;  // one batch at a time
    if (groups[i]->tilt_thr || groups[i]->mem_on)
      return NRC_ERR_UNSUPPORTED;  // tilted fields and fields with memory: single runs only
    for (int j = 0; j < i; ++j)
      if (groups[j] == groups[i]) return NRC_ERR_ARG;
  }
  hipSetDevice(ctx->device);
  nrc_batch* b = new nrc_batch();
  b->ctx = ctx;
  int rc = NRC_OK;
  if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) rc = NRC_ERR_HIP;
  for (int i = 0; rc == NRC_OK && i < n_groups; ++i) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) { rc = NRC_ERR_HIP; break; }
    b->streams.push_back(st);
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) { rc = NRC_ERR_HIP; break; }
    b->done.push_back(ev);
  }
  if (rc != NRC_OK) {
    for (hipStream_t st : b->streams) hipStreamDestroy(st);
    for (hipEvent_t ev : b->done) hipEventDestroy(ev);
    if (b->fork) hipEventDestroy(b->fork);
    delete b;
    return rc;
  }
  hipStreamSynchronize(ctx->stream);
  for (int i = 0; i < n_groups; ++i) {
    b->g.push_back(groups[i]);
    groups[i]->stream = b->streams[i];
  }
  *out = b;
  return NRC_OK;
}

// Runs T slices of every group: configurations config0[i] .. of group i, fields from
// Philox(seed).  graph != 0: the groups' kernels are captured once into one HIP graph (the
// context stream forks to the group streams and joins them again) and launched as a whole;
// the field kernels read seed and config0 from the groups' device parameter blocks, written
// before each launch, so the graph is kept for every later run with the same T.  graph == 0:
// the same kernels are launched directly on the group streams.
extern "C" int nrc_batch_run(nrc_batch* b, uint64_t seed, const uint32_t* config0, int T,
                             int graph) {
  if (!b || !config0 || T < 1) return NRC_ERR_ARG;
  hipSetDevice(b->ctx->device);
  const int n = (int)b->g.size();
  for (nrc_ensemble* e : b->g) {
    e->profile = false;
    NRC_TRY(ens_prepare(e, nullptr, T));
    e->last_T = T;
    e->last_tilted = false;
    e->last_memory = false;
  }
  for (hipStream_t st : b->streams) NRC_HIP(hipStreamSynchronize(st));
  hipStream_t main = b->ctx->stream;
  // seed and configuration ranges go to the groups' parameter blocks on the launching stream,
  // ahead of the fork: the kernels (and a kept graph) read them from there
  for (int i = 0; i < n; ++i) NRC_TRY(ens_set_params(b->g[i], seed, config0[i], main));
  if (!graph) {
    NRC_HIP(hipEventRecord(b->fork, main));
    for (int i = 0; i < n; ++i) {
      NRC_HIP(hipStreamWaitEvent(b->streams[i], b->fork, 0));
      NRC_TRY(ens_enqueue(b->g[i], seed, config0[i], false, T));
      NRC_HIP(hipEventRecord(b->done[i], b->streams[i]));
      NRC_HIP(hipStreamWaitEvent(main, b->done[i], 0));
    }
    NRC_HIP(hipStreamSynchronize(main));
    return NRC_OK;
  }
  if (!(b->exec && b->exec_T == T)) {  // one graph per T: seeds and ranges are parameters
    batch_drop_graph(b);
    hipEvent_t t0, t1;
    hipEventCreate(&t0);
    hipEventCreate(&t1);
    hipEventRecord(t0, main);
    hipEventSynchronize(t0);
    const auto wall0 = std::chrono::steady_clock::now();
    g_capturing = 1;  // before the call: the check build must not synchronise during capture
    if (hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      g_capturing = 0;
      set_detail(__FILE__, __LINE__, "hipStreamBeginCapture failed");
      return NRC_ERR_HIP;
    }
    int rc = NRC_OK;
    if (hipEventRecord(b->fork, main) != hipSuccess) rc = NRC_ERR_HIP;
    for (int i = 0; rc == NRC_OK && i < n; ++i) {
      if (hipStreamWaitEvent(b->streams[i], b->fork, 0) != hipSuccess) { rc = NRC_ERR_HIP; break; }
      rc = ens_enqueue(b->g[i], seed, config0[i], false, T);
      if (rc != NRC_OK) break;
      if (hipEventRecord(b->done[i], b->streams[i]) != hipSuccess ||
          hipStreamWaitEvent(main, b->done[i], 0) != hipSuccess)
        rc = NRC_ERR_HIP;
    }
    hipGraph_t gr = nullptr;
    const hipError_t ce = hipStreamEndCapture(main, &gr);
    g_capturing = 0;
    if (rc == NRC_OK && ce != hipSuccess) rc = NRC_ERR_HIP;
    if (rc == NRC_OK && hipGraphInstantiate(&b->exec, gr, nullptr, nullptr, 0) != hipSuccess)
      rc = NRC_ERR_HIP;
    if (gr) hipGraphDestroy(gr);
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    if (rc != NRC_OK) {
      batch_drop_graph(b);
      return rc;
    }
    b->capture_ms = std::chrono::duration<double, std::milli>(
                        std::chrono::steady_clock::now() - wall0).count();
    b->exec_T = T;
    ++b->captures;
  }
  NRC_HIP(hipGraphLaunch(b->exec, main));
  NRC_HIP(hipStreamSynchronize(main));
  return NRC_OK;
}

extern "C" int nrc_batch_capture_ms(nrc_batch* b, double* ms) {
  if (!b || !ms) return NRC_ERR_ARG;
  *ms = b->capture_ms;
  return NRC_OK;
}

extern "C" int nrc_batch_captures(nrc_batch* b, long* n) {
  if (!b || !n) return NRC_ERR_ARG;
  *n = b->captures;
  return NRC_OK;
}

extern "C" int nrc_batch_destroy(nrc_batch* b) {
  if (!b) return NRC_ERR_ARG;
  hipSetDevice(b->ctx->device);
  for (hipStream_t st : b->streams) hipStreamSynchronize(st);
  hipStreamSynchronize(b->ctx->stream);
  batch_drop_graph(b);
  for (nrc_ensemble* e : b->g) e->stream = b->ctx->stream;
  for (hipStream_t st : b->streams) hipStreamDestroy(st);
  for (hipEvent_t ev : b->done) hipEventDestroy(ev);
  hipEventDestroy(b->fork);
  delete b;
  return NRC_OK;
}

// ------------------------------------------------------------------ slice-by-slice stepping
static bool ens_step_ok(const nrc_ensemble* e) {
  return e->stream == e->ctx->stream && e->d.split == NRC_SPLIT_KSPACE &&
         (e->path == NRC_PATH_SINGLE_PASS ||
          (e->path == NRC_PATH_AUTO && nrc::slice1_supported(e->d.prec, e->d.L)));
}

extern "C" int nrc_ensemble_begin(nrc_ensemble* e, int T) {
  if (!e || T < 1) return NRC_ERR_ARG;
  if (ens_in_batch(e)) return NRC_ERR_BATCHED;
  if (!ens_step_ok(e)) return NRC_ERR_UNSUPPORTED;
  NRC_TRY(ens_prepare(e, nullptr, T));
  e->last_T = T;
  e->last_tilted = false;  // the caller's fields
  e->last_memory = false;
  return ens_set_sources(e);
}

extern "C" int nrc_ensemble_measure(nrc_ensemble* e, int t) {
  if (!e || t < 0 || t >= e->last_T) return NRC_ERR_ARG;
  if (!ens_step_ok(e)) return NRC_ERR_UNSUPPORTED;
  if (e->pending_x) return NRC_ERR_ARG;  // mid-run of evolve_fused: it measures inside
  return ens_measure(e, t);
}

extern "C" int nrc_ensemble_evolve(nrc_ensemble* e, int t, const uint32_t* words_dev) {
  if (!e || !words_dev || t < 0 || t >= e->last_T) return NRC_ERR_ARG;
  if (!ens_step_ok(e)) return NRC_ERR_UNSUPPORTED;
  if (e->pending_x) return NRC_ERR_ARG;
  return ens_evolve(e, t, words_dev);
}

// S slices t .. t + S - 1 in one launch of the persistent kernels with the caller's device
// fields words_dev[S][C][nw], measuring rows t .. t + S - 1 inside (plane-wave sinks): the FP64
// x-chain kernel where the ensemble runs it (between calls the state keeps its x-DFT(+1)
// pending, so measure / evolve are refused until the run ends), else the fused single-pass
// kernel after measuring row t.
extern "C" int nrc_ensemble_evolve_fused(nrc_ensemble* e, int t, int S,
                                         const uint32_t* words_dev) {
  if (!e || !words_dev || S < 1 || S > nrc::SLICE_FUSE_MAX || t < 0 || t + S > e->last_T)
    return NRC_ERR_ARG;
  if (!ens_step_ok(e)) return NRC_ERR_UNSUPPORTED;
  if (e->sinks || e->pair) return NRC_ERR_UNSUPPORTED;  // measured inside: plane waves only
  const size_t slice_words = (size_t)e->d.n_configs * e->nw;
  if (ens_xchain(e, true)) {
    if (t == 0) e->pending_x = false;
    else if (!e->pending_x) return NRC_ERR_ARG;  // a run of the x-chain starts at t = 0
    return ens_evolve_x(e, t, S, words_dev, slice_words, t + S == e->last_T);
  }
  if (e->pending_x) return NRC_ERR_ARG;
  const bool fuse = !e->smear && (!e->trap_tab || nrc::slice1p_trap_ok(e->d.prec, e->d.L));
  if (S > 1 && !fuse) return NRC_ERR_UNSUPPORTED;
  NRC_TRY(ens_measure(e, t));
  if (S > 1) return ens_evolve_fused(e, t, S, words_dev, slice_words);
  return ens_evolve(e, t, words_dev);
}

extern "C" int nrc_ensemble_results_dev(nrc_ensemble* e, const double** logabs_dev,
                                        const double** phase_dev) {
  if (!e || e->last_T < 1) return NRC_ERR_ARG;
  if (logabs_dev) *logabs_dev = e->o_la;
  if (phase_dev) *phase_dev = (const double*)e->o_ph;
  return NRC_OK;
}
