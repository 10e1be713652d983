/* Drives every entry of the FP64 oracle (oracle/oracle.c) on small inputs, for a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitize.py): out-of-bounds
 * accesses, use after free, leaks, signed overflow, misaligned or null accesses abort the
 * run.  Results are not checked here (the oracle's own tests pin them); only that every path
 * runs clean. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double urand(unsigned* s) {
  *s = *s * 1664525u + 1013904223u;
  return (double)(*s >> 8) / 16777216.0 - 0.5;
}

static double* crand(int n, unsigned* s) {
  double* p = (double*)malloc(2 * (size_t)n * sizeof(double));
  for (int i = 0; i < 2 * n; ++i) p[i] = urand(s);
  return p;
}

int main(void) {
  unsigned s = 12345u;
  const int bc0[3] = {0, 0, 0}, bc1[3] = {1, 0, 1};
  /* FFTs: cubic, non-cubic, odd radices, batches */
  const int shapes[][3] = {{4, 4, 4}, {4, 6, 8}, {2, 12, 6}, {5, 3, 7}};
  for (int k = 0; k < 4; ++k) {
    const int V = shapes[k][0] * shapes[k][1] * shapes[k][2];
    double* a = crand(2 * V, &s);
    if (orc_fft3d(a, shapes[k][0], shapes[k][1], shapes[k][2], 2, -1)) return 1;
    if (orc_fft3d(a, shapes[k][0], shapes[k][1], shapes[k][2], 2, +1)) return 1;
    free(a);
  }
  double* b = crand(3 * 15, &s);
  if (orc_fft1d(b, 15, 3, -1)) return 1;
  free(b);
  /* tables */
  const int L = 4, V = L * L * L;
  double* t = (double*)malloc(V * sizeof(double));
  for (int type = 0; type < 3; ++type) {
    orc_dispersion(t, L, L, L, bc1, type, 1.5, 1, M_PI * M_PI);
    orc_kinetic_xi(t, L, L, L, bc0, type, 1.5, 0.5, M_PI * M_PI);
  }
  orc_kinetic_exp(t, L, L, L, 0.05, 1.0);
  if (orc_potential_table(t, L, L, L, 1, 0, 0.1, 0.2, 0.3, 0.05) != 0) return 1;
  if (orc_potential_table(t, L, L, L, 1, 1, 0.1, 0.2, 0.3, 0.05) != 0) return 1;
  const double cpl[3] = {1.0, 0.5, 0.1};
  double* smear = (double*)malloc(V * sizeof(double));
  if (orc_interaction_table(smear, L, L, L, bc0, 5.0, 1, cpl, 3, 1.0, M_PI * M_PI)) return 1;
  double* lat = crand(V, &s);
  double* fld = (double*)malloc(2 * V * sizeof(double));
  if (orc_interaction_field(fld, lat, smear, L, L, L)) return 1;
  /* momenta */
  int moms[3 * 64], par[64];
  double om[64];
  if (orc_momenta(L, L, L, bc0, 2, 1.0, 2.0, M_PI * M_PI, moms, par, om, 64) < 0) return 1;
  int low[3 * 7];
  if (orc_lowest_momenta(L, 7, low)) return 1;
  /* random numbers */
  orc_ranlxs r;
  long st[105];
  for (int lv = 0; lv <= 2; ++lv) {
    if (orc_ranlxs_init(&r, lv, 7012)) return 1;
    for (int i = 0; i < 500; ++i) (void)orc_ranlxs_next(&r);
    orc_ranlxs_get_state(&r, st);
    if (orc_ranlxs_set_state(&r, st)) return 1;
  }
  int8_t phi8[V];
  orc_ranlxs_z2_field(&r, V, phi8);
  double gf[V];
  orc_gauss_field(&r, V, 0.0022, gf);
  (void)orc_gauss(&r, 1.0);
  uint32_t ctr[4] = {1, 2, 3, 4}, key[2] = {5, 6}, out4[4];
  orc_philox4x32_10(ctr, key, out4);
  uint32_t words[3];
  orc_z2_bits(7012, 3, 2, 0, 70, words);
  int8_t phi70[70];
  orc_bits_to_phi(words, 70, phi70);
  /* LU */
  const int n = 6;
  double* a = crand(n * n, &s);
  int indx[6];
  double d = 0, la = 0, ph[2], as = 0;
  if (orc_ludcmp(a, n, indx, &d)) return 1;
  orc_lu_det(a, n, d, &la, ph, &as, NULL);
  free(a);
  /* overlaps, pair sinks, special functions, sources */
  double* p1 = crand(3 * V, &s);
  double* p2 = crand(4 * V, &s);
  double M[2 * 3 * 4];
  orc_overlap(p1, p2, M, 3, 4, V);
  int opp[V];
  orc_opposite_parity_index(L, L, L, bc0, opp);
  double* w = (double*)malloc(V * sizeof(double));
  for (int type = 1; type <= 5; ++type)
    orc_two_body_wavefunc(w, L, L, L, bc0, type, 0, 1, 0.3, 0.2, 1.0, M_PI * M_PI);
  orc_pair_overlap(p1, p2, w, opp, M, 3, 4, V);
  for (int i = -40; i <= 40; ++i) (void)orc_dawson(0.25 * i);
  double h[8];
  orc_hermite(h, 7, 0.3);
  double* coeffs = crand(V, &s);  /* V complex coefficients c(n), n in lattice order */
  double* src = (double*)malloc(2 * V * sizeof(double));
  orc_sho_source(src, coeffs, L, 0.5, 0.6, 0.7);
  /* the correlator path: every field kind, both splits, minors, pair sinks, trap, norms */
  double kin[V];
  orc_kinetic_xi(kin, L, L, L, bc0, 0, 1.0, 0.5, M_PI * M_PI);
  const int T = 3, N = 7;
  int8_t phiT[3 * V];
  for (int i = 0; i < 3 * V; ++i) phiT[i] = (urand(&s) > 0) ? 1 : -1;
  double phiR[3 * V], norms[3] = {1.0, 2.0, 0.5};
  for (int i = 0; i < 3 * V; ++i) phiR[i] = urand(&s);
  double* pot = (double*)malloc(V * sizeof(double));
  if (orc_potential_table(pot, L, L, L, 1, 1, 0.01, 0.01, 0.01, 0.05)) return 1;
  double* sinks = crand(N * V, &s);
  double* sources = crand(N * V, &s);
  double lab[2 * N * 3], phb[4 * N * 3], asb[2 * N * 3];
  double* Mout = (double*)malloc(2 * (size_t)T * N * N * sizeof(double));
  double* psi = (double*)malloc(2 * (size_t)N * V * sizeof(double));
  for (int variant = 0; variant < 8; ++variant) {
    orc_run_desc dsc;
    memset(&dsc, 0, sizeof(dsc));
    dsc.L = L;
    dsc.N = N;
    dsc.T = T;
    dsc.moms = low;
    dsc.kin = kin;
    dsc.A = 0.4;
    dsc.seed = 7012;
    dsc.minors = variant & 1;
    dsc.field_kind = ORC_FIELD_PHILOX;
    orc_ranlxs rr;
    orc_ranlxs_init(&rr, 2, 99);
    switch (variant) {
      case 1: dsc.split = ORC_SPLIT_XSPACE; break;
      case 2: dsc.field_kind = ORC_FIELD_RANLXS; dsc.rng = &rr; break;
      case 3: dsc.field_kind = ORC_FIELD_ARRAY; dsc.phi = phiT; dsc.smear = smear; break;
      case 4: dsc.pair_w = w; dsc.pair_opp = opp; break;
      case 5: dsc.pot = pot; dsc.sources = sources; dsc.sinks = sinks; break;
      case 6: dsc.field_kind = ORC_FIELD_REAL; dsc.phi_real = phiR; dsc.norms = norms; break;
      default: break;
    }
    if (orc_run_config(&dsc, lab, phb, asb, variant == 0 ? Mout : NULL,
                       variant == 0 ? psi : NULL))
      return 1;
  }
  /* the CPU-baseline timer: threads with their own configurations */
  {
    orc_run_desc dsc;
    memset(&dsc, 0, sizeof(dsc));
    dsc.L = L;
    dsc.N = N;
    dsc.T = 2;
    dsc.moms = low;
    dsc.kin = kin;
    dsc.A = 0.4;
    dsc.seed = 1;
    double secs[2];
    if (orc_time_pool(&dsc, 3, 2, 2, 2, secs)) return 1;
    (void)orc_time_threads(&dsc, 2, 2);
  }
  double C[4];
  double ka[8], kb[8];
  for (int i = 0; i < 8; ++i) ka[i] = kb[i] = 1.0 / (1.0 + 0.1 * i);
  if (orc_exact_c1(2, 4, ka, kb, 0.4, C)) return 1;
  free(t); free(smear); free(lat); free(fld); free(p1); free(p2); free(w); free(src);
  free(pot); free(sinks); free(sources); free(Mout); free(psi); free(coeffs);
  puts("oracle clean");
  return 0;
}
