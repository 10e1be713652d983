/*
 * nrcps.h -- C ABI of the MI355X (gfx950) implementation of nrCPS's auxiliary-field
 * propagator and Slater-determinant correlator path.
 *
 * Every entry point returns an int status: NRC_OK (0) or a negative NRC_ERR_* code, and
 * never exits the process (the reference's ERR.General terminates; the C++ class layer in
 * include/nrcps/ turns non-zero codes back into that behaviour).  Pointers named *_dev are
 * device memory obtained from nrc_malloc; everything else is host memory.
 *
 * Data conventions (the reference's): complex values are interleaved (re, im); a volume
 * is [x][y][z] with z fastest (src/utils/vector3.C:175-187); a batch of volumes is
 * contiguous with stride V = L^3 (src/fourier/fourier_common.C:16); sign -1 is
 * FFTW_FORWARD and +1 FFTW_BACKWARD, both unnormalised (src/fourier/noarch/fourier.C:29-35).
 */
#ifndef NRCPS_H
#define NRCPS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NRC_OK 0
#define NRC_ERR_ARG (-1)         /* invalid argument */
#define NRC_ERR_UNSUPPORTED (-2) /* shape or mode this build does not implement */
#define NRC_ERR_HIP (-3)         /* HIP runtime failure */
#define NRC_ERR_NOMEM (-4)       /* device allocation failed */
#define NRC_ERR_SINGULAR (-5)    /* an all-zero matrix row (ludcmp.C:25-28 exits there) */
#define NRC_ERR_BATCHED (-6)     /* the ensemble belongs to a batch (nrc_batch_create) */
#define NRC_ERR_TOO_LARGE (-7)   /* sizes beyond a no-allocation call's fixed scratch */
#define NRC_ERR_CHECK (-8)       /* a check of the NRC_CHECK build failed (bad index, NaN) */

#define NRC_PREC_F32 0
#define NRC_PREC_F64 1

#define NRC_SPLIT_KSPACE 0 /* k-space storage; per slice K(1/2) I K(1/2), field (1+A phi)/V */
#define NRC_SPLIT_XSPACE 1  /* x-space storage; psi <- F+[kappa F-[psi]]/V, then (1 + A phi) */

#define NRC_OVERLAP_F64 0    /* dense overlaps in FP64 */
#define NRC_OVERLAP_F32 1    /* dense overlaps of FP32 data, FP64 accumulation */
#define NRC_OVERLAP_BF16X3 2 /* split-BF16x3 MFMA products, FP32 accumulation */
#define NRC_OVERLAP_BF16X6 3 /* three-way BF16 splits, six MFMA products (FP32-exact inputs),
                                FP32 accumulation */

typedef struct nrc_context nrc_context;
typedef struct nrc_fft nrc_fft;
typedef struct nrc_ensemble nrc_ensemble;

/* ---- context, memory: Cuda::Malloc / Free / MemCopy / DeviceQ (include/cuda_utils.h:10-15) */
int nrc_init(int device, nrc_context** ctx);
int nrc_finalize(nrc_context* ctx);
const char* nrc_error_string(int code);
/* "file:line: what" of the last failure on the calling thread ("" if none yet).  No
 * counterpart in the reference, whose errors print and exit (error.C:36-46). */
const char* nrc_last_error_detail(void);
/* 1 in the check build (libnrcps_hip_check.so, compiled with -DNRC_CHECK), else 0.  The check
 * build, the same ABI, synchronises after every HIP call outside graph capture (a kernel
 * failure is reported by the entry point that launched it), checks index arrays a caller
 * passes (pair_opp sites) before a kernel reads them, and scans FFT outputs, overlap and pair
 * matrices, LU results and ensemble results for NaN (and infinities other than a singular
 * log|det| of -inf), returning NRC_ERR_CHECK with the detail.  For debugging; slower. */
int nrc_check_build(void);
int nrc_device_name(nrc_context* ctx, char* buf, int len, int* n_cu);
int nrc_malloc(nrc_context* ctx, void** p_dev, size_t bytes);
int nrc_free(nrc_context* ctx, void* p_dev);
int nrc_memcpy_h2d(nrc_context* ctx, void* dst_dev, const void* src, size_t bytes);
int nrc_memcpy_d2h(nrc_context* ctx, void* dst, const void* src_dev, size_t bytes);
int nrc_memcpy_d2d(nrc_context* ctx, void* dst_dev, const void* src_dev, size_t bytes);
int nrc_memset(nrc_context* ctx, void* p_dev, int byte, size_t bytes);
int nrc_sync(nrc_context* ctx);

/* ---- FFT: Fourier(batch) / Forward / Backward (include/fourier.h:7-45;
 *      src/fourier/cuda/fourier.cu:9-79).  In-place batched unnormalised 3D C2C (sign -1 is
 *      FFTW_FORWARD) on X x Y x Z lattices with every side in {2, 4, 6, 8, 12, 16, 24, 32,
 *      48, 64}: cubic L <= 32 on two fused passes, other shapes one launch per axis.  Other
 *      sides: NRC_ERR_UNSUPPORTED. */
int nrc_fft_plan(nrc_context* ctx, int lx, int ly, int lz, int batch, int prec, nrc_fft** plan);
int nrc_fft_exec(nrc_fft* plan, void* buf_dev, int sign);
int nrc_fft_destroy(nrc_fft* plan);

/* ---- element-wise operators
 * nrc_scale_real: buf[b][i] *= table[i] (op 0) or /= table[i] (op 1) -- Kinetic::Evolve
 *   (src/hamiltonian/cuda/hamiltonian_evolve.cu:90-104, hamiltonian_kernels.cu:17-42).
 * nrc_mul_complex: buf[b][i] *= field[i] -- the multiply inside Interaction::Evolve
 *   (hamiltonian_evolve.cu:42-66, hamiltonian_kernels.cu:44-57). */
int nrc_scale_real(nrc_context* ctx, void* buf_dev, const void* table_dev, int batch, long vol,
                   int op, int prec);
int nrc_mul_complex(nrc_context* ctx, void* buf_dev, const void* field_dev, int batch, long vol,
                    int prec);
/* nrc_scale_const: buf[i] *= d (op 0) or /= d (op 1) for n complex values -- the loop of
 *   Propagator::Normalize(Float) (src/alg/propagator.C:112-133, host-only in the reference,
 *   NotImplemented in its GPU build).
 * nrc_norms: norms[b] = sqrt(sum |buf[b][i]|^2) of each of batch volumes (FP64 sums), to the
 *   host -- the norms of Propagator::Normalize() (propagator.C:78-110). */
int nrc_scale_const(nrc_context* ctx, void* buf_dev, long n, double d, int op, int prec);
int nrc_norms(nrc_context* ctx, const void* buf_dev, int batch, long vol, int prec,
              double* norms);
/* field = (field + 1)/V in place (re), field/V (im): the end of Interaction::Initialize
 * (src/hamiltonian/cuda/hamiltonian_initialize.cu:30-35, hamiltonian_kernels.cu:59-71). */
int nrc_plus_one_div_vol(nrc_context* ctx, void* buf_dev, long vol, int prec);

/* ---- overlaps: OneBody::Project, Cuda::InnerProduct and the SlaterDet constructor
 *      (src/one_body/cuda/one_body_project.cu:17-34; src/utils/cuda/cuda_utils.cu:88-127;
 *      src/slater_det/slater_det.C:25-39).  M_dev[c][s][t] = sum_i conj(phi[s][i]) psi[c][t][i]
 *      for ncfg batches of nt orbitals (stride vol), complex FP64 output. */
int nrc_overlap(nrc_context* ctx, const void* phi_dev, const void* psi_dev, double* M_dev,
                int ns, int nt, long vol, int ncfg, int mode);
/* nrc_overlap for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The same kernels and results as nrc_overlap, with the split-BF16 modes' workspace (split
 * sinks and per-chunk partial tiles) in a fixed 64 MiB device array that each GPU holds from
 * nrc_init until the process ends; it calls no allocator.  NRC_ERR_TOO_LARGE when the
 * workspace of (ns, nt, vol, ncfg) exceeds 64 MiB (e.g. ns = nt = 33 at vol = 32768 fits
 * about 170 configurations per call).  Calls on one device share the array: calls that
 * overlap in time (other streams, other contexts of the device) race. */
int nrc_overlap_noalloc(nrc_context* ctx, const void* phi_dev, const void* psi_dev,
                        double* M_dev, int ns, int nt, long vol, int ncfg, int mode);
/* GEMM variants of the split-BF16 modes (a plan chosen by measurement, as
 * src/fourier/noarch/fourier.C:30-39 plans its FFTs with FFTW_EXHAUSTIVE):
 *   NRC_OVL_COLS        a real and an imaginary B column per orbital, built in registers,
 *                       four waves per configuration;
 *   NRC_OVL_SHARED      two A rows per sink ((re, im) and (-im, re)), split once per call
 *                       and staged in LDS for the four waves of a workgroup, each wave a
 *                       tile of 48 orbital columns (32 above 56 sinks) taken across
 *                       configurations: 2-step stages, two workgroups per CU;
 *   NRC_OVL_SHARED_DEEP the same with 4-step stages and 4 steps of orbital loads in flight,
 *                       one workgroup per CU;
 *   NRC_OVL_SHARED_PIPE 2-step stages with 4 steps of orbital loads in flight, two
 *                       workgroups per CU.
 * A call uses the context's choice for (ns, nt, vol, mode), else NRC_OVL_SHARED (also for
 * ensembles with dense FP32 sinks).  set_variant pins a choice (NRC_OVL_AUTO clears it);
 * nrc_overlap_tune times every variant (reps timed calls each after an untimed one) on the
 * caller's data, keeps the fastest for the shape and returns it, with each variant's best ms
 * in ms[4] (COLS, SHARED, SHARED_DEEP, SHARED_PIPE order).  The variants differ in FP32 summation order
 * only. */
#define NRC_OVL_AUTO 0
#define NRC_OVL_COLS 1
#define NRC_OVL_SHARED 2
#define NRC_OVL_SHARED_DEEP 3
#define NRC_OVL_SHARED_PIPE 4
int nrc_overlap_set_variant(nrc_context* ctx, int ns, int nt, long vol, int mode, int variant);
int nrc_overlap_variant(nrc_context* ctx, int ns, int nt, long vol, int mode, int* variant);
int nrc_overlap_tune(nrc_context* ctx, const void* phi_dev, const void* psi_dev, double* M_dev,
                     int ns, int nt, long vol, int ncfg, int mode, int reps, int* best,
                     double* ms);

/* ---- the pair-sink matrix: TwoBody::Run(Float*, Float*) and Run(Float**, Float**,
 *      complex<Float>*, int) (include/two_body.h:31-35; src/two_body/noarch/two_body_run.C:
 *      16-43; src/two_body/cuda/two_body_run.cu) for ncfg batches:
 *      M_dev[c][s][t] = sum_i w[i] psi1[c][s][i] psi2[c][t][opp[i]], no conjugation.
 *      psi1: ncfg x n1 x vol, psi2: ncfg x n2 x vol complex (FP64 for NRC_OVERLAP_F64, FP32
 *      otherwise); w: vol doubles (the wave function); opp: vol ints, the site of -(n + bc)
 *      (src/two_body/two_body.C:31-33); all device, w and opp 16-byte aligned for
 *      NRC_OVERLAP_BF16X3.  Complex FP64 output. */
int nrc_pair_overlap(nrc_context* ctx, const void* psi1_dev, const void* psi2_dev,
                     const double* w_dev, const int* opp_dev, double* M_dev, int n1, int n2,
                     long vol, int ncfg, int mode);

/* ---- batched LU log-determinant: ludcmp (include/ludcmp.h:5; src/utils/ludcmp.C:9-65) and
 *      SlaterDet::Run's det / log det (src/slater_det/slater_det.C:59-115).  For each of
 *      ncfg row-major n x n complex matrices (row stride ldm) and, if minors, each leading
 *      minor m = 1..n: log|det|, det/|det| (2 doubles) and the reference's DET_TYPE_LOG
 *      imaginary part.  Outputs [c][m] on the device; status_dev[c][m] = 1 for a zero row. */
int nrc_lu_det(nrc_context* ctx, const double* M_dev, int n, int ldm, int ncfg, int minors,
               double* logabs_dev, double* phase_dev, double* argsum_dev, int* status_dev);

/* ---- ludcmp itself, batched (include/ludcmp.h:5; src/utils/ludcmp.C:9-65): a_dev holds ncfg
 *      contiguous row-major n x n complex matrices, factored in place; indx_dev[c][n] and
 *      d_dev[c] as the reference returns them.  The arithmetic is the reference's, operation
 *      for operation, so pivots and factors are bit-identical to the host build. */
int nrc_ludcmp(nrc_context* ctx, double* a_dev, int n, int ncfg, int* indx_dev, double* d_dev,
               int* status_dev);
/* |z| of n complex values as std::abs(complex<double>) computes it on the host (the modulus
 * ludcmp pivots with, ludcmp.C:24,43). */
int nrc_ref_abs(nrc_context* ctx, const double* z_dev, double* out_dev, long n);

/* ---- auxiliary field: Lattice::Refresh with FIELD_TYPE_ZTWO (src/alg/lattice.C:60-109),
 *      counter-based.  words_dev[c][w], w < ceil(vol/32): Philox4x32-10 block w/4 with
 *      counter (w/4, t, config0 + c, stream), key (seed lo, seed hi); site 32w+j is +1 where
 *      bit j is set, else -1. */
int nrc_z2_bits(nrc_context* ctx, uint64_t seed, uint32_t config0, int nconf, uint32_t t,
                uint32_t stream, int vol, uint32_t* words_dev);

/* ---- ensemble engine: the production slice loop (production/slater_det/main.C:185-220)
 *      for a batch of configurations at once: measure (overlaps of plane-wave sinks, LU),
 *      then evolve one slice (Interaction::Initialize + Propagator::Run with [K, I, K]).
 *      Orbitals, fields and overlap matrices stay on the device. */
typedef struct nrc_ensemble_desc {
  int L;             /* cubic lattice side */
  int n_orb;         /* N orbitals per configuration; sources = sinks (<= 64) */
  int n_configs;     /* configurations evolved together */
  int prec;          /* NRC_PREC_F32 / NRC_PREC_F64 for orbitals and FFTs */
  int split;         /* NRC_SPLIT_KSPACE / NRC_SPLIT_XSPACE */
  int minors;        /* 0: det of the full N x N matrix; 1: every leading minor */
  const int* moms;   /* n_orb x 3 plane-wave momenta (n1, n2, n3) in [0, L) */
  const double* kin; /* V: k-space split xi_half (psi /= xi); x-space split exp(-tau eps) */
  double A;          /* field multiplier (1 + A phi) */
  const double* smear; /* optional V form factor f(k) (k-space split): the field becomes
                          Interaction::Initialize's (1 + F-[f F+ phi])/V */
  const double* sinks; /* optional n_orb x V complex sink vectors in storage space (k-space
                          for the k-space split, x-space for the x-space split).  When set,
                          overlaps are dense products (split-BF16x3 on the matrix cores for
                          FP32, FP64 otherwise) instead of the plane-wave gather; single-pass
                          path only.  The ensemble keeps the sinks and their conjugate mirrors
                          (conj phi(-p) in k-space), 2 n_orb vectors, for conjugate partners.
                          Dense and pair sinks on the FP32 x-chain run several slices per launch
                          and allocate, at the first run, room for the states inside a launch:
                          3 x n_configs x volumes x V complex FP32 (without it, one slice per
                          launch). */
  const double* pair_w; /* optional V real pair wave function w(p) (k-space split, single-pass
                           path).  When set, each measurement is SlaterDet3's
                           (production/slater_det3_trapped/main-cum.C:217-238): the pair
                           matrix phi[s][t] = sum_p w(p) psi_s(p) psi_t(opp(p)) (TwoBody::Run)
                           and 2N results per configuration and slice: m = 2n + 1 the leading
                           (n+1) x (n+1) minor of phi, m = 2n the same with column n replaced by
                           the one-body overlaps <sink_n | orbital_s>. */
  const int* pair_opp;  /* V opposite-parity sites for pair_w (src/two_body/two_body.C:31-33) */
} nrc_ensemble_desc;

int nrc_ensemble_create(nrc_context* ctx, const nrc_ensemble_desc* desc, nrc_ensemble** ens);
/* Fields from Philox(seed) for configs config0 .. config0 + n_configs - 1.  Host outputs
 * (any may be NULL), indexed [t][c][m]: logabs, phase (2 doubles), argsum. */
int nrc_ensemble_run(nrc_ensemble* ens, uint64_t seed, uint32_t config0, int T, double* logabs,
                     double* phase, double* argsum);
/* Same with caller-supplied fields: words[t][c][ceil(V/32)] (host), e.g. from ranlxs. */
int nrc_ensemble_run_fields(nrc_ensemble* ens, const uint32_t* words, int T, double* logabs,
                            double* phase, double* argsum);
/* Orbitals of configuration c after the last run, N x V complex in the ensemble precision,
 * in storage space (k-space for the k-space split).  NRC_ERR_ARG before the first run and
 * after nrc_ensemble_set_symmetry until the next one. */
int nrc_ensemble_orbitals(nrc_ensemble* ens, int c, void* out);
/* Kernel path for the evolution: NRC_PATH_SINGLE_PASS evolves each slice in one kernel that
 * reads and writes every orbital once (FP32 L <= 32, FP64 L <= 16); NRC_PATH_TWO_PASS uses an
 * X pass and a YZ pass per slice (every supported shape); NRC_PATH_AUTO picks per shape. */
#define NRC_PATH_AUTO 0
#define NRC_PATH_TWO_PASS 1
#define NRC_PATH_SINGLE_PASS 2
int nrc_ensemble_set_path(nrc_ensemble* ens, int path);
/* Conjugate partners (conj = 1; 0 turns them off).  Fields are real and the kinetic (and
 * smear) tables even under p -> -p, so the evolution commutes with psi(p) -> conj psi(-p)
 * (k-space; conj psi(x) in x-space) and the orbital of momentum -k stays the conjugate mirror
 * of the one of momentum k.  One orbital of each (k, -k) pair of the set is then evolved and
 * the other's sink values are read conjugated at the mirrored sinks: the same results (to
 * rounding) from (N + self-conjugate orbitals) / 2 volumes per configuration.  The reference
 * evolves every orbital (src/propagator/propagator.C:35-64); this has no counterpart there.
 * Pair sinks: the pair matrices come from two products over the stored volumes (mirrored
 * sites, and the same sites conjugated); they need pair_opp(p) = -p and pair_w even.  Dense
 * sinks: the stored volumes are projected on the sinks and on their conjugate mirrors, since
 * <phi | J psi> = conj <J phi | psi>.  NRC_ERR_UNSUPPORTED with tables (kinetic, smear,
 * pair_w) that are not even to 1e-12 relative; not while the ensemble is in a batch; call
 * before a run. */
int nrc_ensemble_set_symmetry(nrc_ensemble* ens, int conj);
/* Trap (k-space split): every slice the x-space multiplier of the field becomes
 * (1 + A phi) q / V (or the smeared field times q), q = V real values (q = NULL removes the
 * trap).  The reference's trapped slice K, P(1/2), I, P(1/2), K (production/
 * slater_det3_trapped/main-cum.C:97-106; Potential, src/hamiltonian/hamiltonian.C:163-240,
 * noarch/hamiltonian_evolve.C:69-93) is q = (V p)^2 with p Potential's table, since the three
 * x-space multiplies commute.  Not while the ensemble is in a batch. */
int nrc_ensemble_set_trap(nrc_ensemble* ens, const double* q);
/* Arithmetic of FP32 ensembles (k-space split, plain or trapped-free field, plane-wave sinks,
 * single-pass path).  NRC_ACC_FP32: every transform and multiply in FP32 (compensated
 * twiddles and factors).  NRC_ACC_XCHAIN: the state stays FP32 in HBM, registers and LDS,
 * and so do the y and z transforms and the field multiply, but per x-pencil the chain
 * x-DFT(+1) -> sink gather -> kinetic factor -> x-DFT(-1), where FP32 rounding lands on the
 * sink momenta with full weight, runs on FP64 registers rounded once; between launches the
 * state is held with its last x-DFT(+1) pending.  NRC_ACC_X64 (the default): the same with
 * the z transforms of each LDS row also on FP64 registers (except at L = 24, where they
 * would spill).  At L=32, N=33, T=128 the correlator is within ~2e-6 (FP32), ~8e-7 (XCHAIN)
 * and ~2.5e-7 (X64) of FP64; XCHAIN runs ~9% faster than X64, FP32 ~13%.  Other ensembles
 * ignore the setting.  No counterpart in the reference (one Float type everywhere,
 * config.h:19-48).  Not while the ensemble is in a batch. */
#define NRC_ACC_FP32 0
#define NRC_ACC_X64 1
#define NRC_ACC_XCHAIN 2
int nrc_ensemble_set_accuracy(nrc_ensemble* ens, int mode);
/* Importance-sampled Z2 fields (the Z2 form of production/one_body-reweight/main.C:73-76,
 * which draws Lattice::Refresh(phi0), a Gaussian field shifted off zero, and divides the
 * propagator by the density ratio): with p_plus != 0.5, nrc_ensemble_run draws each site +1
 * with probability q = round(p_plus 2^32) / 2^32 (q exactly; site i of configuration c at
 * slice t is +1 where output i % 4 of the Philox4x32-10 block with counter (i / 4, t, c, 2)
 * and key (seed lo, seed hi) is below q 2^32), and keeps each slice's field sum.
 * nrc_ensemble_log_ratio then gives log r[t][c] = sum over t' < t of log(p(phi_t') /
 * q(phi_t')) (p the plain sampler's 2^-V), so that sum_c det(t)^2 r / C estimates C(t)
 * without bias; the ratio is returned beside the results instead of being folded into the
 * orbitals as Propagator::Normalize does.  p_plus = 0.5 restores the plain sampler; caller
 * fields (nrc_ensemble_run_fields) are never tilted (their log ratio is zero).  Not while the
 * ensemble is in a batch, and a batch refuses a tilted ensemble (NRC_ERR_UNSUPPORTED). */
int nrc_ensemble_set_tilt(nrc_ensemble* ens, double p_plus);
int nrc_ensemble_tilt(nrc_ensemble* ens, double* p_plus);  /* the q in use (0.5: plain) */
/* Fields with memory (no counterpart in the reference; an importance sampler for the Z2
 * field of src/lattice/lattice.C:81-86 in the manner of production/one_body-reweight's shifted
 * field): every site remembers its own past field, m(t) = sum_{k >= 1} rho^(k-1) phi(t - k),
 * and at slice t is +1 with probability q = (1 + lambda + theta m) / 2.  Realised exactly in
 * integers: M = 1024 m in int16 with M(t + 1) = floor((r M + 2^14) / 2^15) + 1024 phi(t),
 * r = round(2^15 rho); q = (2^31 + e) / 2^32 with e = clamp(round(2^31 lambda) +
 * round(2^21 theta) M, -2^25, 2^25); site i is +1 where output i % 4 of the Philox4x32-10
 * block with counter (i / 4, t, c, 3) is below 2^31 + e.  |lambda|, |theta| <= 2^-6,
 * 0 <= rho <= 0.96 (NRC_ERR_ARG otherwise); lambda = theta = 0 restores the plain sampler.
 * nrc_ensemble_log_ratio then gives log r[t][c] = sum over slices t' < t and sites of
 * -log1p(phi eps) (eps = e / 2^31), so that sum_c det(t)^2 r / C estimates C(t) without bias.
 * Setting memory turns the tilt off and a tilt turns memory off; not while in a batch, and a
 * batch refuses an ensemble with memory (NRC_ERR_UNSUPPORTED).  nrc_ensemble_memory returns
 * the realised (lambda, theta, rho). */
int nrc_ensemble_set_memory(nrc_ensemble* ens, double lambda, double theta, double rho);
int nrc_ensemble_memory(nrc_ensemble* ens, double* lambda, double* theta, double* rho);
/* The memory network (no counterpart in the reference; the same kind of exact importance
 * sampler): every site carries two memories of its own past field, M_j = 1024 m_j in int16
 * with the fields-with-memory recurrence for rho_0 = 0.75 and rho_1 = 0.94 (r = 24576, 30802),
 * and a pointwise network of 8 ReLU units maps s_j = M_j / 1024 to a logit
 *   a_k = max((W1[k][0] s_0 + W1[k][1] s_1) + b1[k], 0),
 *   l = ((c_0 s_0 + c_1 s_1) + b2 + sum_k W2[k] a_k) 2^-10   (k in order),
 * every operation one FP32 rounding without contraction.  Site i of slice t is +1 where output
 * i % 4 of the Philox4x32-10 block with counter (i / 4, t, c, 4) is below T(l), the learned
 * proposal's FP32 threshold (q(+1) = T / 2^32 exactly); nrc_ensemble_log_ratio then gives
 * log r[t][c] = sum over slices t' < t and sites of log(2^31 / f), f = T or 2^32 - T.
 * params: 35 host floats [W1 (8 x 2), b1 (8), W2 (8), c (2), b2], finite (NRC_ERR_ARG
 * otherwise), or NULL to return to the plain sampler.  Setting it turns the tilt and the
 * fields with memory off (and nrc_ensemble_set_memory / _set_tilt turn it off); not while in
 * a batch, and a batch refuses it like the fields with memory. */
int nrc_ensemble_set_memnet(nrc_ensemble* ens, const float* params);
/* log_ratio: [T][C], T the last run's slice count (NRC_ERR_ARG otherwise) */
int nrc_ensemble_log_ratio(nrc_ensemble* ens, int T, double* log_ratio);
/* Orbital volumes evolved per configuration: n_orb, or fewer with conjugate partners. */
int nrc_ensemble_volumes(nrc_ensemble* ens, int* volumes_per_config);
/* Per-kernel timing with HIP events on the ensemble's stream (enable before a run).
 * kind: 0 X pass or single-pass slice kernel, 1 YZ pass, 2 LU, 3 field generation,
 * 4 sink gather.  Returns total ms and launches. */
int nrc_ensemble_profile(nrc_ensemble* ens, int enable);
int nrc_ensemble_kernel_time(nrc_ensemble* ens, int kind, double* total_ms, long* launches);
/* The per-slice results of the ensemble's last run (nrc_ensemble_run, _run_fields or a batch
 * run), [t][c][m] as nrc_ensemble_run returns them; any output may be NULL. */
int nrc_ensemble_results(nrc_ensemble* ens, double* logabs, double* phase, double* argsum);
/* Autotuner (a plan chosen by measurement, as src/fourier/noarch/fourier.C:30-39 plans with
 * FFTW_EXHAUSTIVE): times over T_probe slices every evolution path the ensemble's shape
 * supports (two-pass FFT passes, single-pass slice kernel); with the path chosen, the slices
 * per launch of the persistent kernels (4, 2, 1; where the path fuses slices, over at least 4
 * slices); for the FP64 x-chain kernel, its tiles (x-pencils per thread, 1 or 2, where the
 * shape offers both); with dense FP32 sinks, the overlap GEMM variants (NRC_OVL_*) on the
 * ensemble's own buffers.  Keeps the fastest of each for later runs and stores the choice in the context
 * under the ensemble's shape key (L, N, precision, split, volumes per configuration,
 * accuracy, measurement kind, configurations), where later ensembles of the same key find it
 * unless nrc_ensemble_set_path chose a path.  Returns the chosen path and its ms per slice.
 * tune_report: the i-th timing of the last tune (kind 0 path, 1 slices per launch, 2 GEMM
 * variant, 3 tile; value; ms per slice, for kind 2 ms per measurement), NRC_ERR_ARG past the
 * end.
 * launch_slices: the slices per launch the ensemble's next run uses. */
int nrc_ensemble_tune(nrc_ensemble* ens, int T_probe, int* path, double* ms_per_slice);
int nrc_ensemble_tune_report(nrc_ensemble* ens, int i, int* kind, int* value, double* ms);
int nrc_ensemble_launch_slices(nrc_ensemble* ens, int* slices);
/* The context's tuned choices (GEMM variants, ensemble paths and slices per launch) to and
 * from a text file, so that a later process starts with them.  A malformed file is refused
 * (NRC_ERR_ARG) and changes nothing; loaded entries replace those of the same key. */
int nrc_tune_save(nrc_context* ctx, const char* path);
int nrc_tune_load(nrc_context* ctx, const char* path);
int nrc_ensemble_destroy(nrc_ensemble* ens);

/* ---- batches of uneven shapes: many configurations of different (L, N) evolved together,
 *      one ensemble per shape group.  No counterpart in the reference, which runs one shape
 *      per process (production/slater_det/main.C:185-247); each group is that loop.
 * nrc_batch_create: the groups (ensembles of this context, in no other batch) get their own
 *   streams; destroy the batch before its ensembles.
 * nrc_batch_run: T slices of every group, group i on configurations config0[i] ..
 *   (Philox(seed) fields).  graph != 0 captures the kernels of all groups and slices into one
 *   HIP graph (fork/join over the group streams) and launches it as one unit; seed and
 *   config0 reach the captured kernels through device parameter blocks, so the graph is
 *   reused by every later run with the same T.  Results: nrc_ensemble_results per group.
 * nrc_batch_capture_ms: host time of the last capture and instantiation.
 * nrc_batch_captures: graphs captured so far (one per distinct T). */
typedef struct nrc_batch nrc_batch;
int nrc_batch_create(nrc_context* ctx, nrc_ensemble* const* groups, int n_groups,
                     nrc_batch** batch);
int nrc_batch_run(nrc_batch* batch, uint64_t seed, const uint32_t* config0, int T, int graph);
int nrc_batch_capture_ms(nrc_batch* batch, double* ms);
int nrc_batch_captures(nrc_batch* batch, long* n);
int nrc_batch_destroy(nrc_batch* batch);

/* ---- slice-by-slice stepping: the loop of nrc_ensemble_run (production/slater_det/
 *      main.C:185-220) split so that the caller chooses each slice's field after seeing the
 *      measurement at t (Lattice::Refresh, lattice.C:60-109, replaced by the caller's
 *      words_dev[c][ceil(V/32)], device memory).  k-space split, single-pass path, on the
 *      context stream (not while the ensemble is in a batch).
 * Stepping runs the plain FP32 kernels for FP32 ensembles (NRC_ACC_FP32 arithmetic).
 * nrc_ensemble_begin: plane-wave sources and result rows for T slices.
 * nrc_ensemble_measure: overlaps and LU at t into result row t.
 * nrc_ensemble_evolve: slice t with the given fields.
 * nrc_ensemble_results_dev: device pointers of the [T][C][m] log|det| and phase rows. */
int nrc_ensemble_begin(nrc_ensemble* ens, int T);
int nrc_ensemble_measure(nrc_ensemble* ens, int t);
int nrc_ensemble_evolve(nrc_ensemble* ens, int t, const uint32_t* words_dev);
/* Slices t .. t + S - 1 (S <= 4) in one launch of the persistent kernels with the caller's
 * fields words_dev[S][n_configs][ceil(V/32)], measuring rows t .. t + S - 1 inside (plane-wave
 * sinks only).  FP32 ensembles at the default accuracy run the FP64 x-chain kernel, whose
 * state keeps an x-transform pending between calls: such a run goes through evolve_fused
 * from t = 0 to the end (measure and evolve are refused in between).  Other ensembles measure
 * row t and run the fused single-pass kernel.  For samplers that draw several slices' fields
 * before evolving them (nrcps_amd/proposal.py LearnedSampler with group > 1). */
int nrc_ensemble_evolve_fused(nrc_ensemble* ens, int t, int S, const uint32_t* words_dev);
int nrc_ensemble_results_dev(nrc_ensemble* ens, const double** logabs_dev,
                             const double** phase_dev);

/* ---- learned field proposal.  No counterpart in the reference, whose fields are always
 *      uniform Z2 (lattice.C:81-86, random.C:129-132); the uniform sampler stays the fallback
 *      and the judge of the proposal's bias.  A 3-layer periodic 3D convolutional network
 *      (3x3x3 kernels, 16 hidden channels, ReLU; BF16 weights and activations, FP32
 *      accumulation) maps phi(t-1) and the configuration's normalised log-weight u to
 *      per-site logits l of P(phi(t) = +1).  Parameters: P = nrc_proposal_param_count()
 *      floats, W1[16][27][2] b1[16] W2[16][27][16] b2[16] W3[27][16] b3, taps 9(dx+1) +
 *      3(dy+1) + (dz+1), out[s] = sum W[tap] in[s + d(tap)] with periodic wrap.
 * nrc_proposal_forward: logits_dev[B][V] of prev_dev[B][nw] (NULL: no previous slice) and
 *   u_dev[B] (all device).
 * nrc_proposal_sample: phi(t) for configurations config0.. into words_dev[B][nw]: phi = +1
 *   where the Philox4x32-10 word (counter (site/4, t, config, stream), key seed; output
 *   site%4) lies below T = ceil(sigmoid(l) 2^32), so q(+1) = T/2^32 exactly, and
 *   logr_dev[c] += sum over sites of log(0.5 / q(phi)) in FP64 (the exact reweighting ratio).
 * nrc_proposal_train_step: gradient of sum_b w_b sum_s -log q(target_b(s)) / (V wsum) into
 *   grad_out (host, may be NULL) and the loss; lr > 0 then takes a step on the FP32 master
 *   parameters: Adam (optimizer 0; 0.9, 0.999, 1e-8) or SGD with momentum 0.9 (optimizer 1).
 * nrc_norm_logw: u_dev[c] = (2 logabs_dev[c*stride] - mu) / sd.
 * nrc_z2_bits_list: the Philox Z2 words of (configs_dev[i], ts_dev[i]) pairs (as nrc_z2_bits). */
typedef struct nrc_proposal nrc_proposal;
int nrc_proposal_create(nrc_context* ctx, int L, int max_batch, nrc_proposal** prop);
int nrc_proposal_destroy(nrc_proposal* prop);
int nrc_proposal_param_count(void);
int nrc_proposal_set_params(nrc_proposal* prop, const float* params);
int nrc_proposal_get_params(nrc_proposal* prop, float* params);
int nrc_proposal_forward(nrc_proposal* prop, const uint32_t* prev_dev, const float* u_dev, int B,
                         float* logits_dev);
int nrc_proposal_sample(nrc_proposal* prop, const uint32_t* prev_dev, const float* u_dev, int B,
                        uint64_t seed, uint32_t config0, uint32_t t, uint32_t stream,
                        uint32_t* words_dev, double* logr_dev);
int nrc_proposal_train_step(nrc_proposal* prop, const uint32_t* prev_dev,
                            const uint32_t* target_dev, const float* u_dev, const float* w_dev,
                            int B, double wsum, float lr, int optimizer, float* grad_out,
                            double* loss_out);
int nrc_norm_logw(nrc_context* ctx, const double* logabs_dev, int stride, int B, double mu,
                  double sd, float* u_dev);
int nrc_z2_bits_list(nrc_context* ctx, uint64_t seed, const uint32_t* configs_dev,
                     const uint32_t* ts_dev, int n, uint32_t stream, int vol,
                     uint32_t* words_dev);
/* sums_dev[t][c] = sum over the vol sites of phi(x, t) for configurations config0 .. config0 +
 * nconf - 1 and t < T, the fields of nrc_z2_bits: the statistic of tilted proposals
 * q ~ p exp(lambda S), whose exact weights are known from S alone. */
int nrc_z2_sums(nrc_context* ctx, uint64_t seed, uint32_t config0, int nconf, int T,
                uint32_t stream, int vol, int* sums_dev);
/* The log ratio log(p / q) that fields with memory (lambda, theta, rho) would carry, evaluated
 * on the plain sampler's fields (stream 0) of configurations config0 .. config0 + nconf - 1:
 * out_dev[t][c] = the sum over slices t' < t (out_dev[0] = 0), for t < T.  With det(t)^2 of
 * the same configurations, E_p[det^2 p / q] is the second moment the memory sampler's estimate
 * would have, which is how its parameters are fitted to a plain sample.  vol <= 32768
 * (NRC_ERR_UNSUPPORTED otherwise). */
int nrc_z2_memory_eval(nrc_context* ctx, uint64_t seed, uint32_t config0, int nconf, int T,
                       int vol, double lambda, double theta, double rho, double* out_dev);
/* The memory network (params, 35 host floats) on the plain sampler's fields of
 * configurations config0 .. config0 + nconf - 1: nrc_z2_memnet_eval writes out_dev[t][c], the
 * cumulative log(p / q) of slices t' < t, as nrc_z2_memory_eval; nrc_z2_memnet_grad writes
 * grad_dev[c][35] = sum_t G_dev[t][c] sum_sites d log q(phi) / d params, with q(+1) =
 * sigmoid(l) in place of T / 2^32 (they agree to ~2e-7), for t < T.  With G[t][c] the
 * weight of slice t of configuration c in the gradient of the effective-sample fraction
 * E_p[det^2]^2 / E_p[det^4 p / q], this trains the network on a plain sample.  Both
 * synchronise; vol <= 32768 (NRC_ERR_UNSUPPORTED otherwise). */
int nrc_z2_memnet_eval(nrc_context* ctx, uint64_t seed, uint32_t config0, int nconf, int T,
                       int vol, const float* params, double* out_dev);
int nrc_z2_memnet_grad(nrc_context* ctx, uint64_t seed, uint32_t config0, int nconf, int T,
                       int vol, const float* params, const double* G_dev, double* grad_dev);

#ifdef __cplusplus
}
#endif
#endif
