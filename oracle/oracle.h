/*
 * oracle.h -- FP64 CPU restatement of nrCPS's auxiliary-field propagator and
 * Slater-determinant correlator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links or calls it.
 *
 * Conventions (all follow the reference):
 *   - complex data is interleaved (re, im) doubles;
 *   - site index i = z + Z*(y + Y*x), z fastest     (src/utils/vector3.C:175-187)
 *   - sign -1 = FFTW_FORWARD ("k -> x"), sign +1 = FFTW_BACKWARD ("x -> k"),
 *     both unnormalised                              (src/fourier/noarch/fourier.C:29-35)
 *   - enum values match include/enum.h.
 *
 * Parity pins for every function are listed in tests/test_oracle.py.
 */
#ifndef NRCPS_ORACLE_H
#define NRCPS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums (values as in include/enum.h of the reference) ---- */
enum { ORC_DISP_STANDARD = 0, ORC_DISP_QUADRATIC = 1, ORC_DISP_PERFECT = 2 };
enum { ORC_CUTOFF_NONE = 0, ORC_CUTOFF_HARD = 1 };
enum { ORC_INTER_NONE = 0, ORC_INTER_XISQMINUSONE = 1, ORC_INTER_ONEMINUSXIINVSQ = 2 };

/* ---- FFT ---- */
/* In-place unnormalised 3D C2C DFT of `batch` contiguous volumes [x][y][z]. */
int orc_fft3d(double* buf, int lx, int ly, int lz, int batch, int sign);
/* Strided 1D DFT of `count` transforms of length n (for tests). */
int orc_fft1d(double* buf, int n, int count, int sign);

/* ---- tables ---- */
/* Dispersion(type, mass, cutoff) (src/alg/dispersion.C:11-73). bc: 0 = PRD, 1 = APRD per axis. */
void orc_dispersion(double* out, int lx, int ly, int lz, const int bc[3], int type,
                    double mass, int cutoff_type, double cutoff_sq);
/* Kinetic xi(k) = 1 + Dispersion(type, mass/dt, HARD)  (src/hamiltonian/hamiltonian.C:260-281). */
void orc_kinetic_xi(double* out, int lx, int ly, int lz, const int bc[3], int type,
                    double mass, double dt, double cutoff_sq);
/* exp(-tau*eps(k)), eps = (1/2M) sum_i 4 sin^2(k_i/2), periodic lattice momenta. */
void orc_kinetic_exp(double* out, int lx, int ly, int lz, double tau, double mass);
/* Interaction form factor f(k) (src/hamiltonian/hamiltonian.C:77-147).  Returns 0, or -1 if
 * the Taylor series is negative somewhere (the reference aborts there). */
/* Potential's table (src/hamiltonian/hamiltonian.C:163-240): v = (k1 x^2 + k2 y^2 + k3 z^2)/2
 * with each coordinate folded to [-L/2, L/2) (form 1 = HARMONIC; form 0 = NONE gives v = 0),
 * then (1 - dt v)/V (type 0 = LIN) or exp(-dt v)/V (type 1 = EXP).  Returns -1 for another
 * form or type. */
enum { ORC_POT_FORM_NONE = 0, ORC_POT_FORM_HARMONIC = 1 };
enum { ORC_POT_TYPE_LIN = 0, ORC_POT_TYPE_EXP = 1 };
int orc_potential_table(double* out, int lx, int ly, int lz, int form, int type, double k1,
                        double k2, double k3, double dt);
int orc_interaction_table(double* out, int lx, int ly, int lz, const int bc[3], double mass,
                          int itype, const double* couplings, int n_couplings, double dt,
                          double cutoff_sq);
/* Interaction::Initialize: fld = (1 + F_-[f * F_+[lat]]) / V, lat interleaved complex
 * (src/hamiltonian/noarch/hamiltonian_initialize.C:14-47). */
int orc_interaction_field(double* fld, const double* lat, const double* f, int lx, int ly, int lz);

/* ---- momenta ---- */
/* Momenta(arg) (src/alg/momenta.C:12-127): momenta with log(1+eps) <= fermi_energy, sorted
 * by energy with the reference's insertion sort.  Writes up to max_count triples into
 * moms (count x 3), the opposite-parity index of each into parity (may be NULL) and the
 * energies into omega (may be NULL).  Returns the count, or -1 if an opposite-parity
 * partner is missing, or -2 if nothing is below the Fermi level. */
int orc_momenta(int lx, int ly, int lz, const int bc[3], int disp_type, double mass,
                double fermi_energy, double cutoff_sq, int* moms, int* parity, double* omega,
                int max_count);
/* The N lowest periodic lattice momenta by |n|^2 (n folded into (-L/2, L/2]); ties keep
 * the enumeration order n1, n2, n3 in [0, L).  Writes N triples in [0,L). */
int orc_lowest_momenta(int L, int n, int* moms);

/* ---- random numbers ---- */
typedef struct {
  int x[96];      /* 96 24-bit numbers: pair k, half h, lane l at 8k + 4h + l */
  int carry[4];
  int pr, prm, ir, jr, is, is_old, level;
} orc_ranlxs;
/* Luscher's ranlxs (luxury level 0/1/2) as in src/random/ranlxs.C:49-206. */
int   orc_ranlxs_init(orc_ranlxs* r, int level, long seed);
float orc_ranlxs_next(orc_ranlxs* r);
void  orc_ranlxs_get_state(const orc_ranlxs* r, long state[105]);
int   orc_ranlxs_set_state(orc_ranlxs* r, const long state[105]);
/* Lattice::Refresh() for FIELD_TYPE_ZTWO: phi_i = 2*floor(2u)-1 for i = 0..V-1 in site
 * order (src/alg/lattice.C:81-86; src/random/random.C:129-132).  Output int8. */
void  orc_ranlxs_z2_field(orc_ranlxs* r, int vol, int8_t* phi);

/* Philox4x32-10 (Salmon et al., SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Counter-based Z2 field of one (config, slice): packed bits, word w of 32 sites
 * (site 32w + j <-> bit j), from Philox block g = w/4 with counter (g, t, config, stream)
 * and key (seed lo, seed hi); word w = out[w % 4].  phi = +1 where the bit is set. */
void orc_z2_bits(uint64_t seed, uint32_t config, uint32_t t, uint32_t stream, int vol,
                 uint32_t* words);
void orc_bits_to_phi(const uint32_t* words, int vol, int8_t* phi);

/* ---- linear algebra ---- */
/* Crout LU with implicit scaled partial pivoting, last maximal row wins (src/utils/ludcmp.C:9-65),
 * on contiguous row-major complex storage.  Returns 0, or -1 for an all-zero row. */
int orc_ludcmp(double* a, int n, int* indx, double* d);
/* From an LU: log|det| = sum log|u_ii|; phase = d * prod u_ii/|u_ii| (unit complex);
 * argsum = arg(d) + sum arg(u_ii) (the imaginary part of the reference's DET_TYPE_LOG);
 * det = d * prod u_ii (DET_TYPE_STANDARD, src/slater_det/slater_det.C:93-109). */
void orc_lu_det(const double* lu, int n, double d, double* logabs, double phase[2],
                double* argsum, double det[2]);
/* M[s][t] = sum_i conj(phi_s[i]) * psi_t[i] (src/one_body/noarch/one_body_project.C:14-27;
 * src/slater_det/slater_det.C:33-37).  phi: ns x vol, psi: nt x vol, M: ns x nt. */
void orc_overlap(const double* phi, const double* psi, double* M, int ns, int nt, long vol);

/* ---- the pair sink (src/two_body/two_body.C:21-143; noarch/two_body_run.C:16-43) ---- */
enum { ORC_WAVEFUNC_NONE = 0, ORC_WAVEFUNC_UNIFORM = 1, ORC_WAVEFUNC_GND = 2,
       ORC_WAVEFUNC_TGND = 3, ORC_WAVEFUNC_PAIR1 = 4, ORC_WAVEFUNC_PAIR2 = 5 };
/* Dawson's integral F(x) = exp(-x^2) int_0^x exp(t^2) dt by Rybicki's sampling-theorem
 * formula with step 0.4 and six odd terms each side, the three-term Taylor series below
 * |x| = 0.2 (src/utils/special_functions.C:48-78; relative accuracy about 2e-7). */
double orc_dawson(double x);
/* out[i] = index of -(n(i) + bc) (src/two_body/two_body.C:31-33). */
void orc_opposite_parity_index(int lx, int ly, int lz, const int bc[3], int* out);
/* TwoBody's wave function w(p) for ORC_WAVEFUNC_* (src/two_body/two_body.C:37-128): GND
 * 1/(lambda - 1/(xi1 xi2)) with xi = 1 + Dispersion(disp, mass, HARD); TGND F(x)/x with
 * x = |p|/(2 lambda), |p|^2 = 2 Dispersion(QUADRATIC, 1, HARD), 0 beyond the cutoff; PAIR1
 * 1/(|p|^2 + lambda); PAIR2 exp(-lambda |p|)/|p|^2 (0 at p = 0); UNIFORM 1/V; NONE 0. */
void orc_two_body_wavefunc(double* out, int lx, int ly, int lz, const int bc[3], int type,
                           int disp1, int disp2, double mass1, double mass2, double lambda,
                           double cutoff_sq);
/* M[s][t] = sum_i w[i] psi1_s[i] psi2_t[opp[i]], no conjugation (TwoBody::Run,
 * src/two_body/noarch/two_body_run.C:16-43, for every pair (s, t) as
 * src/slater_det/slater_det3.C:115-122 fills phi[t + N s]).  M: n1 x n2. */
void orc_pair_overlap(const double* psi1, const double* psi2, const double* w, const int* opp,
                      double* M, int n1, int n2, long vol);

/* ---- SHO sources (src/one_body/one_body.C:53-171; src/utils/special_functions.C:7-34) ---- */
/* h[0..n-1](x): orthonormal Hermite polynomials from h_0 = pi^(-1/4). */
void orc_hermite(double* h, int n, double x);
/* OneBody::Set() for SOURCE_TYPE_SHO on a cubic L^3 lattice: coeffs (V complex) holds c(n) at
 * momentum index n = (n1, n2, n3), n_i < min(10, L); out (V complex) = sum_n c(n) h_n1(d1)
 * h_n2(d2) h_n3(d3) exp(-|d|^2/2) sqrt(l1 l2 l3) sqrt((2 pi)^3 / V) with d = 2 pi k (l1, l2,
 * l2) / L, k centred on 0 (the third axis scaled by l2, as the reference does). */
void orc_sho_source(double* out, const double* coeffs, int L, double l1, double l2, double l3);

/* ---- the correlator path ---- */
enum { ORC_SPLIT_KSPACE = 0,  /* k-space storage, K(1/2) I K(1/2), field (1+A phi)/V in x */
       ORC_SPLIT_XSPACE = 1 }; /* x-space storage, psi <- F+[kappa F-[psi]]/V then (1+A phi) */
enum { ORC_FIELD_PHILOX = 0, ORC_FIELD_RANLXS = 1, ORC_FIELD_ARRAY = 2, ORC_FIELD_REAL = 3 };

typedef struct {
  int L;               /* cubic lattice side */
  int N;               /* orbitals (= sinks = sources) per spin */
  int T;               /* time slices */
  int split;           /* ORC_SPLIT_* */
  int minors;          /* 0: det of the full N x N matrix; 1: all leading minors n = 1..N */
  const int* moms;     /* N momenta (n1, n2, n3) in [0, L): sources and sinks */
  const double* kin;   /* V real: 1/xi_half (k-space split) or exp(-tau eps) (x-space split) */
  double A;            /* field strength: multiplier (1 + A phi) */
  const double* smear; /* optional V real form factor f(k); when set, the reference field is
                          Interaction::Initialize's (1 + F-[f F+ phi])/V and A is unused */
  int field_kind;      /* ORC_FIELD_* */
  uint64_t seed;       /* PHILOX */
  uint32_t config;     /* PHILOX config index */
  orc_ranlxs* rng;     /* RANLXS: consumed sequentially, T*V draws */
  const int8_t* phi;   /* ARRAY: T x V values in {-1, +1} */
  const double* pair_w; /* optional V real pair wave function (k-space split only): the
                           measurement is SlaterDet3's (production/slater_det3_trapped/
                           main-cum.C:217-238), 2N determinants per slice */
  const int* pair_opp;  /* V opposite-parity indices for pair_w */
  const double* sources; /* optional N x V complex sources = sinks in storage space (k-space
                            for the k-space split) instead of the plane waves of moms */
  const double* pot;     /* optional V real Potential table (orc_potential_table, with its
                            1/V; k-space split): each slice is then the trapped production
                            order K, P, I, P, K (production/slater_det3_trapped/main-cum.C:
                            97-106), every P its own forward / backward transform pair */
  const double* phi_real; /* REAL: T x V real field values (e.g. shifted Gaussians) */
  const double* sinks;    /* optional N x V complex sinks in storage space (default: the
                             sources) */
  const double* norms;    /* optional T values: after slice t the orbitals are divided by
                             norms[t] (Propagator::Normalize(Float), src/alg/propagator.C:
                             116-133) */
  int gather;             /* 1, with plane-wave sinks and sources in the k-space split: each
                             overlap <k_s|psi_t> is read from psi_t at k_s instead of the
                             N^2 full-length products of the reference (one_body_project.C:
                             20-23); the values are the same, the products adding exact
                             zeros.  Not the baseline: a timing variant */
} orc_run_desc;

/* Random::Gauss(sigma) (src/random/random.C:105-116): the polar Box-Muller method on the
 * ranlxs uniforms (x, y from 2u - 1, rejected while r2 > 1 or r2 == 0), returning only
 * sigma y sqrt(-2 log r2 / r2).  orc_gauss_field: Lattice::Refresh for FIELD_TYPE_GAUSS
 * (src/alg/lattice.C:60-109) plus a real shift (Refresh(Float), lattice.C:111-165): out[i] =
 * Gauss(1) + shift for the V sites in order (the imaginary parts draw nothing). */
double orc_gauss(orc_ranlxs* r, double sigma);
void orc_gauss_field(orc_ranlxs* r, int vol, double shift, double* out);

/* Runs one configuration through T slices, measuring before evolving at each t
 * (production/slater_det/main.C:185-220).  Outputs, for each t and each measured minor m
 * (m = 0 only, or m = 0..N-1 for minor size m+1; with pair_w m = 0..2N-1: at 2n + 1 the
 * leading (n+1) x (n+1) minor of the pair matrix phi[s][t], at 2n the same minor with
 * column n replaced by the one-body overlaps <sink_n | orbital_s>, SlaterDet3::Run(n+1, -1,
 * -1) and Run(n+1, -1, n), src/slater_det/slater_det3.C:140-210), at [t*nm + m]:
 *   logabs, phase (2 doubles), argsum.  M_out (optional): T x N x N complex overlaps.
 * psi_out (optional): final N x V orbitals in storage space. */
int orc_run_config(const orc_run_desc* d, double* logabs, double* phase, double* argsum,
                   double* M_out, double* psi_out);

/* Exact E[S(t)^2] for a single k=0 orbital with iid Z2 fields:
 * g_0 = 1/V, g <- F^-1[kb^2 F[(1 + A^2 delta_r0) F^-1[ka^2 F g]]], C(t) = sum_r g_t(r).
 * ka / kb are the kinetic multipliers before / after the field (V real, even in k). */
int orc_exact_c1(int L, int T, const double* ka, const double* kb, double A, double* C);

/* Wall time of `nthreads` independent configuration streams, each running `slices` slices
 * of configuration `d` (config index offset per thread), in seconds. */
double orc_time_threads(const orc_run_desc* d, int nthreads, int slices);

/* The CPU-baseline protocol: `nthreads` streams, each warmed up by one slice of its own
 * configuration, then `reps` barrier-synchronised repetitions of `configs` configurations x
 * `slices` slices; secs[r] = wall seconds of repetition r.  Returns 0, -1 on failure. */
int orc_time_pool(const orc_run_desc* d, int nthreads, int configs, int slices, int reps,
                  double* secs);

#ifdef __cplusplus
}
#endif
#endif
