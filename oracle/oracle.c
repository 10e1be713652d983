/*
 * oracle.c -- FP64 CPU restatement of the nrCPS auxiliary-field propagator and
 * Slater-determinant correlator path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP product path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * Every routine names the reference file:line it restates.  The LU, the dispersion tables
 * and the field multiplies keep the reference's expression order so that results agree
 * to the last bit where the reference's own code is deterministic (the LU is checked
 * bit-for-bit against the reference's ludcmp.C compiled into oracle/_ref).  The FFT is
 * this file's own mixed-radix Stockham transform; FFTW (which the reference calls) is
 * not available, so FFT results are pinned by closed forms, a brute-force DFT and numpy.
 *
 * Parity pins (tests/test_oracle.py): ranlxs vs Luscher's published vectors
 * (tests/ranlx/main.C); ludcmp vs tests/ludcmp_test/main.C's known answer, the Cauchy
 * determinant and oracle/_ref; FFT vs brute-force DFT / numpy; dispersion and interaction
 * tables vs their closed forms; Philox vs Random123's published vectors; the correlator
 * run vs free-gas closed forms, the exact Z2 expectation and brute force at L=2.
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* constants as written in include/constants.h:12-14 */
static const double REF_PI = 3.1415926535897932;
static const double REF_TWOPI = 6.2831853071795864;
static const double REF_PISQ = 9.8696044010893586;

/* =========================================================================================
 * FFT: mixed-radix Stockham autosort (radix 4, 2, then odd primes), FP64.
 * ========================================================================================= */
#define ORC_MAXR 64

typedef struct {
  int n, nf;
  int fac[40];
  double* tw; /* e^{-2 pi i k / n}, k < n, interleaved */
} orc_plan1d;

static int orc_plan_init(orc_plan1d* p, int n) {
  int m = n, nf = 0;
  if (n < 1) return -1;
  while (m % 4 == 0) { p->fac[nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[nf++] = 2; m /= 2; }
  for (int q = 3; m > 1; q += 2)
    while (m % q == 0) {
      if (q > ORC_MAXR) return -1;
      p->fac[nf++] = q;
      m /= q;
    }
  p->n = n;
  p->nf = nf;
  p->tw = (double*)malloc(2 * (size_t)n * sizeof(double));
  if (!p->tw) return -1;
  for (int k = 0; k < n; ++k) {
    double a = 2.0 * M_PI * (double)k / (double)n;
    p->tw[2 * k] = cos(a);
    p->tw[2 * k + 1] = -sin(a);
  }
  return 0;
}

static void orc_plan_free(orc_plan1d* p) { free(p->tw); p->tw = NULL; }

/* out[m] = sum_r v[r] w^{r m}, w = e^{sign 2 pi i / R}; tw is the length-n table. */
static void orc_small_dft(int R, const double* v, double* out, int n, const double* tw,
                          int sign) {
  if (R == 2) {
    out[0] = v[0] + v[2]; out[1] = v[1] + v[3];
    out[2] = v[0] - v[2]; out[3] = v[1] - v[3];
    return;
  }
  if (R == 4) {
    double s02r = v[0] + v[4], s02i = v[1] + v[5];
    double d02r = v[0] - v[4], d02i = v[1] - v[5];
    double s13r = v[2] + v[6], s13i = v[3] + v[7];
    double d13r = v[2] - v[6], d13i = v[3] - v[7];
    /* sign -1: out1 = d02 - i d13 ; sign +1: out1 = d02 + i d13 */
    double jr = (sign < 0) ? d13i : -d13i, ji = (sign < 0) ? -d13r : d13r;
    out[0] = s02r + s13r; out[1] = s02i + s13i;
    out[2] = d02r + jr;   out[3] = d02i + ji;
    out[4] = s02r - s13r; out[5] = s02i - s13i;
    out[6] = d02r - jr;   out[7] = d02i - ji;
    return;
  }
  int step = n / R;
  for (int m = 0; m < R; ++m) {
    double ar = 0.0, ai = 0.0;
    for (int r = 0; r < R; ++r) {
      int e = ((r * m) % R) * step;
      double c = tw[2 * e], s = (sign < 0) ? tw[2 * e + 1] : -tw[2 * e + 1];
      ar += v[2 * r] * c - v[2 * r + 1] * s;
      ai += v[2 * r] * s + v[2 * r + 1] * c;
    }
    out[2 * m] = ar;
    out[2 * m + 1] = ai;
  }
}

/* One length-n transform of contiguous data a (n complex); b is scratch of equal size.
 * Returns the buffer holding the result. */
static double* orc_fft_line(const orc_plan1d* p, double* a, double* b, int sign) {
  const int n = p->n;
  double v[2 * ORC_MAXR], t[2 * ORC_MAXR];
  double *in = a, *out = b;
  int Ns = 1;
  for (int f = 0; f < p->nf; ++f) {
    const int R = p->fac[f];
    const int nR = n / R;
    const int tstep = n / (Ns * R);
    for (int j = 0; j < nR; ++j) {
      const int k = j % Ns;
      for (int r = 0; r < R; ++r) {
        double re = in[2 * (j + r * nR)], im = in[2 * (j + r * nR) + 1];
        int e = (r * k * tstep) % n;
        if (e == 0) {
          v[2 * r] = re; v[2 * r + 1] = im;
        } else {
          double c = p->tw[2 * e], s = (sign < 0) ? p->tw[2 * e + 1] : -p->tw[2 * e + 1];
          v[2 * r] = re * c - im * s;
          v[2 * r + 1] = re * s + im * c;
        }
      }
      orc_small_dft(R, v, t, n, p->tw, sign);
      const int d = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; ++r) {
        out[2 * (d + r * Ns)] = t[2 * r];
        out[2 * (d + r * Ns) + 1] = t[2 * r + 1];
      }
    }
    double* s = in; in = out; out = s;
    Ns *= R;
  }
  return in;
}

/* transform `count` lines of length n: line l starts at base + l*... given by the caller */
static void orc_fft_axis(double* buf, const orc_plan1d* p, long stride, long nlines,
                         const long* starts, int sign, double* w1, double* w2) {
  const int n = p->n;
  for (long l = 0; l < nlines; ++l) {
    double* s = buf + 2 * starts[l];
    for (int i = 0; i < n; ++i) {
      w1[2 * i] = s[2 * i * stride];
      w1[2 * i + 1] = s[2 * i * stride + 1];
    }
    double* r = orc_fft_line(p, w1, w2, sign);
    for (int i = 0; i < n; ++i) {
      s[2 * i * stride] = r[2 * i];
      s[2 * i * stride + 1] = r[2 * i + 1];
    }
  }
}

int orc_fft1d(double* buf, int n, int count, int sign) {
  orc_plan1d p;
  if (orc_plan_init(&p, n)) return -1;
  double* w = (double*)malloc(4 * (size_t)n * sizeof(double));
  long* starts = (long*)malloc((size_t)count * sizeof(long));
  for (int c = 0; c < count; ++c) starts[c] = (long)c * n;
  orc_fft_axis(buf, &p, 1, count, starts, sign, w, w + 2 * n);
  free(starts);
  free(w);
  orc_plan_free(&p);
  return 0;
}

int orc_fft3d(double* buf, int lx, int ly, int lz, int batch, int sign) {
  orc_plan1d px, py, pz;
  if (orc_plan_init(&px, lx)) return -1;
  if (orc_plan_init(&py, ly)) { orc_plan_free(&px); return -1; }
  if (orc_plan_init(&pz, lz)) { orc_plan_free(&px); orc_plan_free(&py); return -1; }
  const long V = (long)lx * ly * lz;
  int nmax = lx > ly ? lx : ly;
  if (lz > nmax) nmax = lz;
  double* w = (double*)malloc(4 * (size_t)nmax * sizeof(double));
  long nl = (long)ly * lz;
  if ((long)lx * lz > nl) nl = (long)lx * lz;
  if ((long)lx * ly > nl) nl = (long)lx * ly;
  long* starts = (long*)malloc((size_t)nl * sizeof(long));
  for (int b = 0; b < batch; ++b) {
    const long off = (long)b * V;
    long c = 0;
    for (int x = 0; x < lx; ++x)
      for (int y = 0; y < ly; ++y) starts[c++] = off + ((long)x * ly + y) * lz;
    orc_fft_axis(buf, &pz, 1, c, starts, sign, w, w + 2 * nmax);
    c = 0;
    for (int x = 0; x < lx; ++x)
      for (int z = 0; z < lz; ++z) starts[c++] = off + (long)x * ly * lz + z;
    orc_fft_axis(buf, &py, lz, c, starts, sign, w, w + 2 * nmax);
    c = 0;
    for (int y = 0; y < ly; ++y)
      for (int z = 0; z < lz; ++z) starts[c++] = off + (long)y * lz + z;
    orc_fft_axis(buf, &px, (long)ly * lz, c, starts, sign, w, w + 2 * nmax);
  }
  free(starts);
  free(w);
  orc_plan_free(&px);
  orc_plan_free(&py);
  orc_plan_free(&pz);
  return 0;
}

/* =========================================================================================
 * Tables
 * ========================================================================================= */

/* src/alg/dispersion.C:11-73, expression order kept. */
void orc_dispersion(double* out, int lx, int ly, int lz, const int bc[3], int type,
                    double mass, int cutoff_type, double cutoff_sq) {
  double sh[3];
  for (int a = 0; a < 3; ++a) sh[a] = bc[a] ? 0.5 : 0.0;
  long i = 0;
  for (int x0 = 0; x0 < lx; ++x0)
    for (int y0 = 0; y0 < ly; ++y0)
      for (int z0 = 0; z0 < lz; ++z0, ++i) {
        int x = x0, y = y0, z = z0;
        if (type == ORC_DISP_STANDARD) {
          double sx = sin((x + sh[0]) * REF_PI / lx);
          double sy = sin((y + sh[1]) * REF_PI / ly);
          double sz = sin((z + sh[2]) * REF_PI / lz);
          double e = 2.0 * (sx * sx + sy * sy + sz * sz);
          e /= mass;
          out[i] = e;
        } else {
          if (x >= lx / 2) x = lx - x - bc[0];
          if (y >= ly / 2) y = ly - y - bc[1];
          if (z >= lz / 2) z = lz - z - bc[2];
          double sx = (x + sh[0]) / (double)lx;
          double sy = (y + sh[1]) / (double)ly;
          double sz = (z + sh[2]) / (double)lz;
          double psq = (sx * sx + sy * sy + sz * sz) * 4.0 * REF_PISQ;
          if (cutoff_type == ORC_CUTOFF_HARD && psq > cutoff_sq) {
            out[i] = INFINITY;
          } else if (type == ORC_DISP_QUADRATIC) {
            out[i] = psq / (2.0 * mass);
          } else {
            out[i] = (exp(psq / (2.0 * mass)) - 1.0);
          }
        }
      }
}

/* src/hamiltonian/hamiltonian.C:260-281 */
void orc_kinetic_xi(double* out, int lx, int ly, int lz, const int bc[3], int type,
                    double mass, double dt, double cutoff_sq) {
  const long V = (long)lx * ly * lz;
  orc_dispersion(out, lx, ly, lz, bc, type, mass / dt, ORC_CUTOFF_HARD, cutoff_sq);
  for (long i = 0; i < V; ++i) out[i] = 1.0 + out[i];
}

void orc_kinetic_exp(double* out, int lx, int ly, int lz, double tau, double mass) {
  long i = 0;
  for (int x = 0; x < lx; ++x)
    for (int y = 0; y < ly; ++y)
      for (int z = 0; z < lz; ++z, ++i) {
        double sx = sin(M_PI * x / lx), sy = sin(M_PI * y / ly), sz = sin(M_PI * z / lz);
        double eps = (4.0 * (sx * sx + sy * sy + sz * sz)) / (2.0 * mass);
        out[i] = exp(-tau * eps);
      }
}

/* src/hamiltonian/hamiltonian.C:163-240 (Potential::Potential); the site index is
 * z + Z (y + Y x) and each coordinate n >= L/2 becomes n - L */
int orc_potential_table(double* out, int lx, int ly, int lz, int form, int type, double k1,
                        double k2, double k3, double dt) {
  if (form != ORC_POT_FORM_NONE && form != ORC_POT_FORM_HARMONIC) return -1;
  if (type != ORC_POT_TYPE_LIN && type != ORC_POT_TYPE_EXP) return -1;
  const long V = (long)lx * ly * lz;
  for (long i = 0; i < V; ++i) {
    int z = (int)(i % lz), y = (int)((i / lz) % ly), x = (int)(i / ((long)lz * ly));
    if (x >= lx / 2) x -= lx;
    if (y >= ly / 2) y -= ly;
    if (z >= lz / 2) z -= lz;
    double v = 0.0;
    if (form == ORC_POT_FORM_HARMONIC) {
      v = k1 * x * x;
      v += k2 * y * y;
      v += k3 * z * z;
      v /= 2.0;
    }
    const double p = type == ORC_POT_TYPE_LIN ? 1.0 - dt * v : exp(-dt * v);
    out[i] = p / (double)V;
  }
  return 0;
}

/* Potential::Evolve (src/hamiltonian/noarch/hamiltonian_evolve.C:69-93) on N volumes:
 * Forward, multiply by the table, Backward */
static void orc_potential_evolve(double* psi, const double* pot, int L, int N) {
  const long V = (long)L * L * L;
  orc_fft3d(psi, L, L, L, N, -1);
  for (int n = 0; n < N; ++n) {
    double* a = psi + 2 * (size_t)n * V;
    for (long i = 0; i < V; ++i) { a[2 * i] *= pot[i]; a[2 * i + 1] *= pot[i]; }
  }
  orc_fft3d(psi, L, L, L, N, +1);
}

/* src/hamiltonian/hamiltonian.C:77-147 */
int orc_interaction_table(double* out, int lx, int ly, int lz, const int bc[3], double mass,
                          int itype, const double* couplings, int n_couplings, double dt,
                          double cutoff_sq) {
  const long V = (long)lx * ly * lz;
  orc_dispersion(out, lx, ly, lz, bc, ORC_DISP_PERFECT, mass, ORC_CUTOFF_NONE, cutoff_sq);
  const double XI = exp(cutoff_sq / (2.0 * mass));
  int bad = 0;
  for (long i = 0; i < V; ++i) {
    double xi = 1.0 + out[i];
    if (xi > XI) xi = XI;
    double psq;
    if (itype == ORC_INTER_ONEMINUSXIINVSQ) psq = 1.0 - 1.0 / (xi * xi);
    else if (itype == ORC_INTER_XISQMINUSONE) psq = xi * xi - 1.0;
    else return -2;
    psq *= mass;
    double v = 0.0;
    for (int j = 0; j < n_couplings; ++j) v += couplings[j] * pow(psq, j);
    if (v < 0.0) bad = 1;
    v *= 2.0 * REF_TWOPI / mass;
    v = sqrt(v);
    v /= V;
    v *= dt;
    out[i] = v;
  }
  return bad ? -1 : 0;
}

/* src/hamiltonian/noarch/hamiltonian_initialize.C:14-47 */
int orc_interaction_field(double* fld, const double* lat, const double* f, int lx, int ly,
                          int lz) {
  const long V = (long)lx * ly * lz;
  memcpy(fld, lat, 2 * (size_t)V * sizeof(double));
  if (orc_fft3d(fld, lx, ly, lz, 1, +1)) return -1; /* field->Backward() */
  for (long i = 0; i < V; ++i) {
    fld[2 * i] *= f[i];
    fld[2 * i + 1] *= f[i];
  }
  if (orc_fft3d(fld, lx, ly, lz, 1, -1)) return -1; /* field->Forward() */
  for (long i = 0; i < V; ++i) {
    fld[2 * i] += 1.0;
    fld[2 * i] /= V;
    fld[2 * i + 1] /= V;
  }
  return 0;
}

/* =========================================================================================
 * Momenta
 * ========================================================================================= */

/* src/alg/momenta.C:12-127 */
int orc_momenta(int lx, int ly, int lz, const int bc[3], int disp_type, double mass,
                double fermi_energy, double cutoff_sq, int* moms, int* parity, double* omega,
                int max_count) {
  const long V = (long)lx * ly * lz;
  double* disp = (double*)malloc((size_t)V * sizeof(double));
  orc_dispersion(disp, lx, ly, lz, bc, disp_type, mass, ORC_CUTOFF_HARD, cutoff_sq);
  int count = 0;
  for (int n1 = 0; n1 < lx; ++n1)
    for (int n2 = 0; n2 < ly; ++n2)
      for (int n3 = 0; n3 < lz; ++n3) {
        double e = log(1 + disp[n3 + n2 * lz + n1 * ly * lz]);
        if (e <= fermi_energy) ++count;
      }
  if (count == 0) { free(disp); return -2; }
  double* om = (double*)malloc((size_t)count * sizeof(double));
  int* mm = (int*)malloc(3 * (size_t)count * sizeof(int));
  int c = 0;
  for (int n1 = 0; n1 < lx; ++n1)
    for (int n2 = 0; n2 < ly; ++n2)
      for (int n3 = 0; n3 < lz; ++n3) {
        double e = log(1 + disp[n3 + n2 * lz + n1 * ly * lz]);
        if (e <= fermi_energy) {
          om[c] = e;
          mm[3 * c] = n1; mm[3 * c + 1] = n2; mm[3 * c + 2] = n3;
          ++c;
        }
      }
  free(disp);
  /* insertion sort exactly as the reference, including its `j > 0` bound */
  for (int i = 1; i < count; ++i) {
    double e = om[i];
    int t0 = mm[3 * i], t1 = mm[3 * i + 1], t2 = mm[3 * i + 2];
    int j = i - 1;
    while (j > 0 && om[j] > e) {
      om[j + 1] = om[j];
      mm[3 * (j + 1)] = mm[3 * j]; mm[3 * (j + 1) + 1] = mm[3 * j + 1];
      mm[3 * (j + 1) + 2] = mm[3 * j + 2];
      --j;
    }
    om[j + 1] = e;
    mm[3 * (j + 1)] = t0; mm[3 * (j + 1) + 1] = t1; mm[3 * (j + 1) + 2] = t2;
  }
  int rc = count;
  const int dims[3] = {lx, ly, lz};
  for (int i = 0; i < count && rc >= 0; ++i) {
    int p[3];
    for (int a = 0; a < 3; ++a) p[a] = (dims[a] - mm[3 * i + a] - bc[a]) % dims[a];
    int found = -1;
    for (int j = 0; j < count; ++j)
      if (mm[3 * j] == p[0] && mm[3 * j + 1] == p[1] && mm[3 * j + 2] == p[2]) found = j;
    if (found < 0) rc = -1;
    else if (parity && i < max_count) parity[i] = found;
  }
  for (int i = 0; i < count && i < max_count; ++i) {
    if (moms) { moms[3 * i] = mm[3 * i]; moms[3 * i + 1] = mm[3 * i + 1]; moms[3 * i + 2] = mm[3 * i + 2]; }
    if (omega) omega[i] = om[i];
  }
  free(om);
  free(mm);
  return rc;
}

int orc_lowest_momenta(int L, int n, int* moms) {
  int got = 0;
  for (int q = 0; got < n && q <= 3 * L * L; ++q)
    for (int a = 0; a < L && got < n; ++a)
      for (int b = 0; b < L && got < n; ++b)
        for (int c = 0; c < L && got < n; ++c) {
          int fa = a > L / 2 ? a - L : a, fb = b > L / 2 ? b - L : b, fc = c > L / 2 ? c - L : c;
          if (fa * fa + fb * fb + fc * fc == q) {
            moms[3 * got] = a; moms[3 * got + 1] = b; moms[3 * got + 2] = c;
            ++got;
          }
        }
  return got == n ? 0 : -1;
}

/* =========================================================================================
 * ranlxs: the 24-bit subtract-with-borrow generator with Luscher's luxury decimation,
 * four interleaved lanes, in the state layout of src/random/ranlxs.C:49-206 so that the
 * output sequence and the 105-long state vector are identical.
 * ========================================================================================= */
#define RLX_BASE 0x1000000
#define RLX_MASK 0xffffff

static int rlx_pr_of_level(int level) {
  return level == 0 ? 109 : level == 1 ? 202 : level == 2 ? 397 : -1;
}

int orc_ranlxs_init(orc_ranlxs* r, int level, long seed) {
  int pr = rlx_pr_of_level(level);
  if (pr < 0) return -1;
  if (seed <= 0 || seed > 2147483647L) return -2;
  int bits[31];
  long s = seed;
  for (int k = 0; k < 31; ++k) { bits[k] = (int)(s % 2); s /= 2; }
  int ib = 0, jb = 18;
  for (int lane = 0; lane < 4; ++lane)
    for (int k = 0; k < 24; ++k) {
      int v = 0;
      for (int l = 0; l < 24; ++l) {
        v = 2 * v + bits[ib];
        bits[ib] = (bits[ib] + bits[jb]) % 2;
        ib = (ib + 1) % 31;
        jb = (jb + 1) % 31;
      }
      if (k % 4 == lane) v = RLX_MASK - v;
      r->x[4 * k + lane] = v;
    }
  for (int l = 0; l < 4; ++l) r->carry[l] = 0;
  r->level = level;
  r->pr = pr;
  r->prm = pr % 12;
  r->ir = 0;
  r->jr = 7;
  r->is = 95;
  r->is_old = 0;
  return 0;
}

static void rlx_update(orc_ranlxs* r) {
  int i = r->ir, j = r->jr;
  for (int step = 0; step < r->pr; ++step) {
    int* xi = &r->x[8 * i];
    const int* xj = &r->x[8 * j];
    for (int l = 0; l < 4; ++l) {
      int d = xj[l] - xi[l] - r->carry[l];
      if (d < 0) { xi[4 + l] += 1; d += RLX_BASE; }
      xi[l] = d & RLX_MASK;
    }
    for (int l = 0; l < 4; ++l) {
      int d = xj[4 + l] - xi[4 + l];
      r->carry[l] = d < 0;
      d += RLX_BASE;
      xi[4 + l] = d & RLX_MASK;
    }
    i = (i + 1) % 12;
    j = (j + 1) % 12;
  }
  r->ir = (r->ir + r->prm) % 12;
  r->jr = (r->jr + r->prm) % 12;
  r->is = 8 * r->ir;
  r->is_old = r->is;
}

float orc_ranlxs_next(orc_ranlxs* r) {
  r->is = (r->is + 1) % 96;
  if (r->is == r->is_old) rlx_update(r);
  return (float)ldexp(1.0, -24) * (float)r->x[r->is];
}

void orc_ranlxs_get_state(const orc_ranlxs* r, long state[105]) {
  state[0] = 105;
  for (int k = 0; k < 96; ++k) state[k + 1] = r->x[k];
  for (int l = 0; l < 4; ++l) state[97 + l] = r->carry[l];
  state[101] = r->pr;
  state[102] = r->ir;
  state[103] = r->jr;
  state[104] = r->is;
}

int orc_ranlxs_set_state(orc_ranlxs* r, const long state[105]) {
  if (state[0] != 105) return -1;
  for (int k = 0; k < 96; ++k)
    if (state[k + 1] < 0 || state[k + 1] > RLX_MASK) return -1;
  for (int l = 0; l < 4; ++l)
    if (state[97 + l] != 0 && state[97 + l] != 1) return -1;
  long pr = state[101], ir = state[102], jr = state[103], is = state[104];
  if ((pr != 109 && pr != 202 && pr != 397) || ir < 0 || ir > 11 || jr != (ir + 7) % 12 ||
      is < 0 || is > 95)
    return -1;
  for (int k = 0; k < 96; ++k) r->x[k] = (int)state[k + 1];
  for (int l = 0; l < 4; ++l) r->carry[l] = (int)state[97 + l];
  r->pr = (int)pr;
  r->prm = (int)(pr % 12);
  r->level = pr == 109 ? 0 : pr == 202 ? 1 : 2;
  r->ir = (int)ir;
  r->jr = (int)jr;
  r->is = (int)is;
  r->is_old = 8 * r->ir;
  return 0;
}

void orc_ranlxs_z2_field(orc_ranlxs* r, int vol, int8_t* phi) {
  for (int i = 0; i < vol; ++i) {
    double u = (double)orc_ranlxs_next(r);
    int z = (int)floor(2 * u);
    phi[i] = (int8_t)(2 * z - 1);
  }
}

/* =========================================================================================
 * Philox4x32-10 and the counter-based Z2 field
 * ========================================================================================= */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    if (round) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_z2_bits(uint64_t seed, uint32_t config, uint32_t t, uint32_t stream, int vol,
                 uint32_t* words) {
  const int nw = (vol + 31) / 32;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int g = 0; 4 * g < nw; ++g) {
    uint32_t ctr[4] = {(uint32_t)g, t, config, stream}, out[4];
    orc_philox4x32_10(ctr, key, out);
    for (int q = 0; q < 4 && 4 * g + q < nw; ++q) words[4 * g + q] = out[q];
  }
}

void orc_bits_to_phi(const uint32_t* words, int vol, int8_t* phi) {
  for (int i = 0; i < vol; ++i) phi[i] = ((words[i >> 5] >> (i & 31)) & 1u) ? 1 : -1;
}

/* =========================================================================================
 * LU (src/utils/ludcmp.C:9-65 on contiguous storage) and determinants
 * ========================================================================================= */
#define ORC_TINY 1.0e-20

int orc_ludcmp(double* ad, int n, int* indx, double* d) {
  double complex* a = (double complex*)ad;
  double* vv = (double*)malloc((size_t)n * sizeof(double));
  int imax = 0;
  *d = 1.0;
  for (int i = 0; i < n; ++i) {
    double big = 0.0, temp;
    for (int j = 0; j < n; ++j)
      if ((temp = cabs(a[n * i + j])) > big) big = temp;
    if (big == 0.0) { free(vv); return -1; }
    vv[i] = 1.0 / big;
  }
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < j; ++i) {
      double complex sum = a[n * i + j];
      for (int k = 0; k < i; ++k) sum -= a[n * i + k] * a[n * k + j];
      a[n * i + j] = sum;
    }
    double big = 0.0;
    for (int i = j; i < n; ++i) {
      double complex sum = a[n * i + j];
      for (int k = 0; k < j; ++k) sum -= a[n * i + k] * a[n * k + j];
      a[n * i + j] = sum;
      double dum = vv[i] * cabs(sum);
      if (dum >= big) { big = dum; imax = i; }
    }
    if (j != imax) {
      for (int k = 0; k < n; ++k) {
        double complex t = a[n * imax + k];
        a[n * imax + k] = a[n * j + k];
        a[n * j + k] = t;
      }
      *d = -(*d);
      vv[imax] = vv[j];
    }
    indx[j] = imax;
    if (cabs(a[n * j + j]) == 0.0) a[n * j + j] = ORC_TINY;
    if (j != n) {
      const double complex one = CMPLX(1.0, 0.0);
      double complex dum = one / a[n * j + j];
      for (int i = j + 1; i < n; ++i) a[n * i + j] *= dum;
    }
  }
  free(vv);
  return 0;
}

void orc_lu_det(const double* lu, int n, double d, double* logabs, double phase[2],
                double* argsum, double det[2]) {
  const double complex* a = (const double complex*)lu;
  double la = 0.0;
  double complex ph = CMPLX(d, 0.0);
  double complex dt = CMPLX(d, 0.0);
  double complex ld = clog(CMPLX(d, 0.0));
  for (int i = 0; i < n; ++i) {
    double complex u = a[n * i + i];
    double m = cabs(u);
    la += log(m);
    ph *= u / m;
    dt *= u;
    ld += clog(u);
  }
  if (logabs) *logabs = la;
  if (phase) {
    double m = cabs(ph);
    phase[0] = creal(ph) / m;
    phase[1] = cimag(ph) / m;
  }
  if (argsum) *argsum = cimag(ld);
  if (det) { det[0] = creal(dt); det[1] = cimag(dt); }
}

void orc_overlap(const double* phi, const double* psi, double* M, int ns, int nt, long vol) {
  for (int s = 0; s < ns; ++s)
    for (int t = 0; t < nt; ++t) {
      const double* a = phi + 2 * (size_t)s * vol;
      const double* b = psi + 2 * (size_t)t * vol;
      double re = 0.0, im = 0.0;
      for (long i = 0; i < 2 * vol; i += 2) {
        re += a[i] * b[i] + a[i + 1] * b[i + 1];
        im += a[i] * b[i + 1] - a[i + 1] * b[i];
      }
      M[2 * ((size_t)s * nt + t)] = re;
      M[2 * ((size_t)s * nt + t) + 1] = im;
    }
}

/* =========================================================================================
 * The pair sink
 * ========================================================================================= */
/* Rybicki (1989): F(x) = lim_{h->0} pi^-1/2 sum over odd n of exp(-(x - n h)^2) / n.  With
 * n0 the even integer nearest x/h and x' = x - n0 h, the odd n = n0 +- (2i - 1) give
 * exp(-x'^2) c_i (e^{2 x' (2i-1) h} / (n0 + 2i - 1) + e^{-2 x' (2i-1) h} / (n0 - 2i + 1))
 * with c_i = exp(-((2i - 1) h)^2); the exponentials are stepped by e^{4 x' h}.
 * Below 0.2: x (1 - 2/3 x^2 (1 - 2/5 x^2 (1 - 2/7 x^2))), the Taylor series to x^7.
 * Step, term count and evaluation order are special_functions.C:48-78's, so the values
 * agree to rounding. */
double orc_dawson(double x) {
  const double h = 0.4;
  if (fabs(x) < 0.2) {
    const double x2 = x * x;
    return x * (1.0 - (2.0 / 3.0) * x2 * (1.0 - 0.4 * x2 * (1.0 - (2.0 / 7.0) * x2)));
  }
  double c[7];
  for (int i = 1; i <= 6; ++i) c[i] = exp(-((2.0 * i - 1.0) * h) * ((2.0 * i - 1.0) * h));
  const double ax = fabs(x);
  const int n0 = 2 * (int)(0.5 * ax / h + 0.5);
  const double xp = ax - n0 * h;
  double e1 = exp(2.0 * xp * h);
  const double e2 = e1 * e1;
  double d1 = n0 + 1, d2 = d1 - 2.0, sum = 0.0;
  for (int i = 1; i <= 6; ++i, d1 += 2.0, d2 -= 2.0, e1 *= e2)
    sum += c[i] * (e1 / d1 + 1.0 / (d2 * e1));
  const double m = exp(-xp * xp);
  return (x >= 0.0 ? m : -m) * sum / 1.7724538509055160273;
}

static int wrap_mod(int a, int n) { return ((a % n) + n) % n; }

void orc_opposite_parity_index(int lx, int ly, int lz, const int bc[3], int* out) {
  for (int x = 0; x < lx; ++x)
    for (int y = 0; y < ly; ++y)
      for (int z = 0; z < lz; ++z)
        out[((long)x * ly + y) * lz + z] =
            (wrap_mod(-(x + bc[0]), lx) * ly + wrap_mod(-(y + bc[1]), ly)) * lz +
            wrap_mod(-(z + bc[2]), lz);
}

void orc_two_body_wavefunc(double* out, int lx, int ly, int lz, const int bc[3], int type,
                           int disp1, int disp2, double mass1, double mass2, double lambda,
                           double cutoff_sq) {
  const long V = (long)lx * ly * lz;
  double* e1 = (double*)malloc((size_t)V * sizeof(double));
  double* e2 = (double*)malloc((size_t)V * sizeof(double));
  if (type == ORC_WAVEFUNC_GND) {
    orc_dispersion(e1, lx, ly, lz, bc, disp1, mass1, ORC_CUTOFF_HARD, cutoff_sq);
    orc_dispersion(e2, lx, ly, lz, bc, disp2, mass2, ORC_CUTOFF_HARD, cutoff_sq);
  } else if (type >= ORC_WAVEFUNC_TGND) {
    orc_dispersion(e1, lx, ly, lz, bc, ORC_DISP_QUADRATIC, 1.0, ORC_CUTOFF_HARD, cutoff_sq);
  }
  for (long i = 0; i < V; ++i) {
    double w = 0.0;
    if (type == ORC_WAVEFUNC_UNIFORM) {
      w = 1.0 / (double)V;
    } else if (type == ORC_WAVEFUNC_GND) {
      const double xi1 = 1.0 + e1[i], xi2 = 1.0 + e2[i];
      w = 1.0;
      w /= lambda - 1.0 / (xi1 * xi2);
    } else if (type == ORC_WAVEFUNC_TGND) {
      const double psq = 2.0 * e1[i];
      if (psq > cutoff_sq) {
        w = 0.0;
      } else {
        const double xi = sqrt(psq) / (2.0 * lambda);
        w = xi < 1e-15 ? 1.0 : orc_dawson(xi) / xi;
      }
    } else if (type == ORC_WAVEFUNC_PAIR1) {
      w = 1.0 / (2.0 * e1[i] + lambda);
    } else if (type == ORC_WAVEFUNC_PAIR2) {
      const double psq = 2.0 * e1[i];
      w = psq < 1e-15 ? 0.0 : exp(-lambda * sqrt(psq)) / psq;
    }
    out[i] = w;
  }
  free(e1);
  free(e2);
}

void orc_pair_overlap(const double* psi1, const double* psi2, const double* w, const int* opp,
                      double* M, int n1, int n2, long vol) {
  for (int s = 0; s < n1; ++s)
    for (int t = 0; t < n2; ++t) {
      const double* a = psi1 + 2 * (size_t)s * vol;
      const double* b = psi2 + 2 * (size_t)t * vol;
      double re = 0.0, im = 0.0;
      for (long i = 0; i < vol; ++i) {
        const long k = opp[i];
        re += (a[2 * i] * b[2 * k] - a[2 * i + 1] * b[2 * k + 1]) * w[i];
        im += (a[2 * i] * b[2 * k + 1] + a[2 * i + 1] * b[2 * k]) * w[i];
      }
      M[2 * ((size_t)s * n2 + t)] = re;
      M[2 * ((size_t)s * n2 + t) + 1] = im;
    }
}

/* =========================================================================================
 * SHO sources
 * ========================================================================================= */
void orc_hermite(double* h, int n, double x) {
  double a = 0.0;
  h[0] = 0.7511255444649425; /* pi^(-1/4) */
  for (int j = 1; j < n; ++j) {
    const double b = a;
    a = h[j - 1];
    h[j] = x * sqrt(2.0 / j) * a - sqrt((double)(j - 1) / j) * b;
  }
}

void orc_sho_source(double* out, const double* coeffs, int L, double l1, double l2, double l3) {
  const long V = (long)L * L * L;
  const int nm = L < 10 ? L : 10;
  const double tp = 6.2831853071795864;
  const double norm = sqrt(l1 * l2 * l3) * sqrt(tp * tp * tp / (double)V);
  double h1[10], h2[10], h3[10];
  for (long i = 0; i < V; ++i) {
    int x = (int)(i / ((long)L * L)), y = (int)((i / L) % L), z = (int)(i % L);
    if (x >= L / 2) x -= L;
    if (y >= L / 2) y -= L;
    if (z >= L / 2) z -= L;
    const double dx = tp * x * l1 / L, dy = tp * y * l2 / L, dz = tp * z * l2 / L;
    orc_hermite(h1, nm, dx);
    orc_hermite(h2, nm, dy);
    orc_hermite(h3, nm, dz);
    double re = 0.0, im = 0.0;
    for (int a = 0; a < nm; ++a)
      for (int b = 0; b < nm; ++b)
        for (int d = 0; d < nm; ++d) {
          const long j = ((long)a * L + b) * L + d;
          const double t = h1[a] * h2[b] * h3[d];
          re += coeffs[2 * j] * t;
          im += coeffs[2 * j + 1] * t;
        }
    const double g = exp(-(dx * dx + dy * dy + dz * dz) / 2.0) * norm;
    out[2 * i] = re * g;
    out[2 * i + 1] = im * g;
  }
}

/* =========================================================================================
 * The correlator run (production/slater_det/main.C:185-220 for one configuration)
 * ========================================================================================= */
static void orc_plane_wave(double* out, int L, const int k[3], int xspace) {
  const long V = (long)L * L * L;
  memset(out, 0, 2 * (size_t)V * sizeof(double));
  if (!xspace) { /* OneBody::Set(kx, ky, kz): unit vector in k-space (one_body.C:215-228) */
    out[2 * (((long)k[0] * L + k[1]) * L + k[2])] = 1.0;
    return;
  }
  const double nrm = 1.0 / sqrt((double)V);
  long i = 0;
  for (int x = 0; x < L; ++x)
    for (int y = 0; y < L; ++y)
      for (int z = 0; z < L; ++z, ++i) {
        long ph = ((long)k[0] * x + (long)k[1] * y + (long)k[2] * z) % L;
        double a = 2.0 * M_PI * (double)ph / (double)L;
        out[2 * i] = cos(a) * nrm;
        out[2 * i + 1] = sin(a) * nrm;
      }
}

double orc_gauss(orc_ranlxs* r, double sigma) {
  double x, y, r2;
  do {
    x = 2.0 * (double)orc_ranlxs_next(r) - 1.0;
    y = 2.0 * (double)orc_ranlxs_next(r) - 1.0;
    r2 = x * x + y * y;
  } while (r2 > 1.0 || r2 == 0.0);
  return sigma * y * sqrt(-2.0 * log(r2) / r2);
}

void orc_gauss_field(orc_ranlxs* r, int vol, double shift, double* out) {
  for (int i = 0; i < vol; ++i) out[i] = orc_gauss(r, 1.0) + shift;
}

static void orc_get_field(const orc_run_desc* d, int t, long V, int8_t* phi, uint32_t* words) {
  if (d->field_kind == ORC_FIELD_PHILOX) {
    orc_z2_bits(d->seed, d->config, (uint32_t)t, 0u, (int)V, words);
    orc_bits_to_phi(words, (int)V, phi);
  } else if (d->field_kind == ORC_FIELD_RANLXS) {
    orc_ranlxs_z2_field(d->rng, (int)V, phi);
  } else if (d->field_kind == ORC_FIELD_ARRAY) {
    memcpy(phi, d->phi + (size_t)t * V, (size_t)V);
  }
}

/* the real field value of site i at slice t: the +-1 of phi, or the REAL array's value */
static double orc_field_value(const orc_run_desc* d, int t, long V, const int8_t* phi, long i) {
  return d->field_kind == ORC_FIELD_REAL ? d->phi_real[(size_t)t * V + i] : (double)phi[i];
}

int orc_run_config(const orc_run_desc* d, double* logabs, double* phase, double* argsum,
                   double* M_out, double* psi_out) {
  const int L = d->L, N = d->N, T = d->T;
  const long V = (long)L * L * L;
  const int xspace = d->split == ORC_SPLIT_XSPACE;
  const int pair = d->pair_w != NULL;
  const int nm = pair ? 2 * N : (d->minors ? N : 1);
  if (pair && (xspace || !d->pair_opp)) return -1;
  if (d->pot && xspace) return -1;
  double* psi = (double*)malloc(2 * (size_t)N * V * sizeof(double));
  double* sink = (double*)malloc(2 * (size_t)N * V * sizeof(double));
  double* fld = (double*)malloc(2 * (size_t)V * sizeof(double));
  double* lat = (double*)malloc(2 * (size_t)V * sizeof(double));
  double* kv = (double*)malloc((size_t)V * sizeof(double));
  double* M = (double*)malloc(2 * (size_t)N * N * sizeof(double));
  double* Mp = (double*)malloc(2 * (size_t)N * N * sizeof(double));
  double* sub = (double*)malloc(2 * (size_t)N * N * sizeof(double));
  int* indx = (int*)malloc((size_t)N * sizeof(int));
  int8_t* phi = (int8_t*)malloc((size_t)V);
  uint32_t* words = (uint32_t*)malloc(((size_t)V / 32 + 4) * sizeof(uint32_t));
  int rc = 0;
  if (!psi || !sink || !fld || !lat || !kv || !M || !Mp || !sub || !indx || !phi || !words) {
    rc = -3;
    goto done;
  }
  for (int n = 0; n < N; ++n) {
    if (d->sources)
      memcpy(psi + 2 * (size_t)n * V, d->sources + 2 * (size_t)n * V,
             2 * (size_t)V * sizeof(double));
    else
      orc_plane_wave(psi + 2 * (size_t)n * V, L, d->moms + 3 * n, xspace);
    memcpy(sink + 2 * (size_t)n * V, d->sinks ? d->sinks + 2 * (size_t)n * V
                                              : psi + 2 * (size_t)n * V,
           2 * (size_t)V * sizeof(double));
  }
  if (xspace)
    for (long i = 0; i < V; ++i) kv[i] = d->kin[i] / (double)V;

  for (int t = 0; t < T; ++t) {
    /* measure: SlaterDet ctor + Run(n) */
    if (d->gather && !xspace && !d->sources && !d->sinks) {
      for (int s = 0; s < N; ++s) {
        const int* k = d->moms + 3 * s;
        const size_t js = ((size_t)k[0] * L + k[1]) * L + k[2];
        for (int q = 0; q < N; ++q) {
          M[2 * (s * N + q)] = psi[2 * ((size_t)q * V + js)];
          M[2 * (s * N + q) + 1] = psi[2 * ((size_t)q * V + js) + 1];
        }
      }
    } else {
      orc_overlap(sink, psi, M, N, N, V);
    }
    if (pair) orc_pair_overlap(psi, psi, d->pair_w, d->pair_opp, Mp, N, N, V);
    if (M_out)
      memcpy(M_out + 2 * (size_t)t * N * N, pair ? Mp : M, 2 * (size_t)N * N * sizeof(double));
    for (int m = 0; m < nm; ++m) {
      const int n = pair ? m / 2 + 1 : (d->minors ? m + 1 : N);
      const double* src = pair ? Mp : M;
      for (int s = 0; s < n; ++s)
        for (int q = 0; q < n; ++q) {
          sub[2 * (s * n + q)] = src[2 * (s * N + q)];
          sub[2 * (s * n + q) + 1] = src[2 * (s * N + q) + 1];
        }
      if (pair && (m & 1) == 0) /* column n - 1: psi1[s + (n-1) N] = <sink_{n-1}|orbital_s> */
        for (int s = 0; s < n; ++s) {
          sub[2 * (s * n + n - 1)] = M[2 * ((n - 1) * N + s)];
          sub[2 * (s * n + n - 1) + 1] = M[2 * ((n - 1) * N + s) + 1];
        }
      double dd, la = -INFINITY, ph[2] = {0.0, 0.0}, as = 0.0;
      if (orc_ludcmp(sub, n, indx, &dd) == 0) orc_lu_det(sub, n, dd, &la, ph, &as, NULL);
      const size_t o = (size_t)t * nm + m;
      if (logabs) logabs[o] = la;
      if (phase) { phase[2 * o] = ph[0]; phase[2 * o + 1] = ph[1]; }
      if (argsum) argsum[o] = as;
    }
    /* evolve with the field of this slice */
    orc_get_field(d, t, V, phi, words);
    if (!xspace) {
      if (d->smear) {
        for (long i = 0; i < V; ++i) {
          lat[2 * i] = orc_field_value(d, t, V, phi, i);
          lat[2 * i + 1] = 0.0;
        }
        orc_interaction_field(fld, lat, d->smear, L, L, L);
      } else {
        for (long i = 0; i < V; ++i) {
          fld[2 * i] = (1.0 + d->A * orc_field_value(d, t, V, phi, i)) / (double)V;
          fld[2 * i + 1] = 0.0;
        }
      }
      for (int n = 0; n < N; ++n) { /* Kinetic::Evolve, dt = 1/2 */
        double* a = psi + 2 * (size_t)n * V;
        for (long i = 0; i < V; ++i) { a[2 * i] /= d->kin[i]; a[2 * i + 1] /= d->kin[i]; }
      }
      if (d->pot) orc_potential_evolve(psi, d->pot, L, N);
      orc_fft3d(psi, L, L, L, N, -1); /* Interaction::Evolve: Forward */
      for (int n = 0; n < N; ++n) {
        double* a = psi + 2 * (size_t)n * V;
        for (long i = 0; i < V; ++i) {
          double re = fld[2 * i] * a[2 * i] - fld[2 * i + 1] * a[2 * i + 1];
          double im = fld[2 * i] * a[2 * i + 1] + fld[2 * i + 1] * a[2 * i];
          a[2 * i] = re;
          a[2 * i + 1] = im;
        }
      }
      orc_fft3d(psi, L, L, L, N, +1); /* Backward */
      if (d->pot) orc_potential_evolve(psi, d->pot, L, N);
      for (int n = 0; n < N; ++n) {
        double* a = psi + 2 * (size_t)n * V;
        for (long i = 0; i < V; ++i) { a[2 * i] /= d->kin[i]; a[2 * i + 1] /= d->kin[i]; }
      }
    } else {
      orc_fft3d(psi, L, L, L, N, -1);
      for (int n = 0; n < N; ++n) {
        double* a = psi + 2 * (size_t)n * V;
        for (long i = 0; i < V; ++i) { a[2 * i] *= kv[i]; a[2 * i + 1] *= kv[i]; }
      }
      orc_fft3d(psi, L, L, L, N, +1);
      for (int n = 0; n < N; ++n) {
        double* a = psi + 2 * (size_t)n * V;
        for (long i = 0; i < V; ++i) {
          double f = 1.0 + d->A * orc_field_value(d, t, V, phi, i);
          a[2 * i] *= f;
          a[2 * i + 1] *= f;
        }
      }
    }
    if (d->norms) /* Propagator::Normalize(norm): every value divided by norm */
      for (size_t i = 0; i < 2 * (size_t)N * V; ++i) psi[i] /= d->norms[t];
  }
  if (psi_out) memcpy(psi_out, psi, 2 * (size_t)N * V * sizeof(double));
done:
  free(psi); free(sink); free(fld); free(lat); free(kv); free(M); free(Mp); free(sub); free(indx);
  free(phi); free(words);
  return rc;
}

int orc_exact_c1(int L, int T, const double* ka, const double* kb, double A, double* C) {
  const long V = (long)L * L * L;
  double* g = (double*)calloc(2 * (size_t)V, sizeof(double));
  if (!g) return -1;
  for (long i = 0; i < V; ++i) g[2 * i] = 1.0 / (double)V;
  for (int t = 0; t < T; ++t) {
    double s = 0.0;
    for (long i = 0; i < V; ++i) s += g[2 * i];
    C[t] = s;
    if (t == T - 1) break;
    orc_fft3d(g, L, L, L, 1, -1);
    for (long i = 0; i < V; ++i) { double f = ka[i] * ka[i] / (double)V; g[2 * i] *= f; g[2 * i + 1] *= f; }
    orc_fft3d(g, L, L, L, 1, +1);
    g[0] *= 1.0 + A * A;
    g[1] *= 1.0 + A * A;
    orc_fft3d(g, L, L, L, 1, -1);
    for (long i = 0; i < V; ++i) { double f = kb[i] * kb[i] / (double)V; g[2 * i] *= f; g[2 * i + 1] *= f; }
    orc_fft3d(g, L, L, L, 1, +1);
  }
  free(g);
  return 0;
}

/* =========================================================================================
 * Threaded timing of independent configuration streams (one per host core, as the
 * reference's MPI ranks run independent streams, src/random/random.C:88)
 * ========================================================================================= */
typedef struct {
  orc_run_desc d;
  int rc;
} orc_thread_arg;

static void* orc_thread_main(void* p) {
  orc_thread_arg* a = (orc_thread_arg*)p;
  a->rc = orc_run_config(&a->d, NULL, NULL, NULL, NULL, NULL);
  return NULL;
}

/* The CPU baseline's timing protocol: nthreads independent streams (one per core), each
 * first runs a warm-up configuration of one slice, then `reps` repetitions of `configs`
 * configurations x `slices` slices; every repetition starts on a barrier across the threads
 * and its wall time is taken from the first thread's start to the last thread's end.
 * secs[r] receives repetition r's wall seconds.  Returns 0, or -1 on failure. */
typedef struct {
  orc_run_desc d;
  int configs, reps, rc;
  pthread_barrier_t* bar;
  struct timespec* t_start;  /* per repetition, written by thread 0 after the barrier */
} orc_pool_arg;

static void* orc_pool_main(void* p) {
  orc_pool_arg* a = (orc_pool_arg*)p;
  orc_run_desc w = a->d;
  w.T = 1;
  w.config = a->d.config + 0x40000000u;
  a->rc |= orc_run_config(&w, NULL, NULL, NULL, NULL, NULL);
  for (int r = 0; r < a->reps; ++r) {
    pthread_barrier_wait(a->bar);
    if (a->t_start) clock_gettime(CLOCK_MONOTONIC, &a->t_start[r]);
    for (int c = 0; c < a->configs; ++c) {
      orc_run_desc d = a->d;
      d.config = a->d.config + (uint32_t)(r * a->configs + c) * 4096u;
      a->rc |= orc_run_config(&d, NULL, NULL, NULL, NULL, NULL);
    }
    pthread_barrier_wait(a->bar);
  }
  return NULL;
}

int orc_time_pool(const orc_run_desc* d, int nthreads, int configs, int slices, int reps,
                  double* secs) {
  if (nthreads < 1 || configs < 1 || slices < 1 || reps < 1 || d->field_kind == ORC_FIELD_RANLXS)
    return -1;
  orc_pool_arg* args = (orc_pool_arg*)calloc((size_t)nthreads, sizeof(orc_pool_arg));
  pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
  struct timespec* t0 = (struct timespec*)calloc((size_t)reps, sizeof(struct timespec));
  pthread_barrier_t bar;
  pthread_barrier_init(&bar, NULL, (unsigned)nthreads + 1);
  for (int i = 0; i < nthreads; ++i) {
    args[i].d = *d;
    args[i].d.T = slices;
    args[i].d.config = d->config + (uint32_t)i;
    args[i].configs = configs;
    args[i].reps = reps;
    args[i].bar = &bar;
    args[i].t_start = i == 0 ? t0 : NULL;
  }
  for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, orc_pool_main, &args[i]);
  for (int r = 0; r < reps; ++r) {
    pthread_barrier_wait(&bar); /* every thread warmed up / done with repetition r - 1 */
    pthread_barrier_wait(&bar); /* every thread done with repetition r */
    struct timespec t1;
    clock_gettime(CLOCK_MONOTONIC, &t1);
    secs[r] = (double)(t1.tv_sec - t0[r].tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0[r].tv_nsec);
  }
  for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
  int bad = 0;
  for (int i = 0; i < nthreads; ++i) bad |= args[i].rc;
  pthread_barrier_destroy(&bar);
  free(args);
  free(th);
  free(t0);
  return bad ? -1 : 0;
}

double orc_time_threads(const orc_run_desc* d, int nthreads, int slices) {
  if (nthreads < 1 || d->field_kind == ORC_FIELD_RANLXS) return -1.0;
  orc_thread_arg* args = (orc_thread_arg*)calloc((size_t)nthreads, sizeof(orc_thread_arg));
  pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
  for (int i = 0; i < nthreads; ++i) {
    args[i].d = *d;
    args[i].d.T = slices;
    args[i].d.config = d->config + (uint32_t)i;
  }
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, orc_thread_main, &args[i]);
  for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  int bad = 0;
  for (int i = 0; i < nthreads; ++i) bad |= args[i].rc;
  free(args);
  free(th);
  if (bad) return -1.0;
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
