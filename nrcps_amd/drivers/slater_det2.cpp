// slater_det2.cpp -- the two-species pair-sink correlator driver on the MI355X class layer.
//
// The measurement loop of production/slater_det2_untrapped/main-cum.C:110-400 with the
// reference's class API and file formats: momentum sources for the N = Momenta count plane
// waves below the Fermi energy (one_body_arg forced to SOURCE_TYPE_MOM and momenta_arg's
// dispersion and mass to kinetic_arg's, as main-cum.C:78-96 does), the species-2 orbitals
// being the species-1 orbitals reindexed by OppositeParityIndex (main-cum.C:240-245), the
// pair wave function of args/two_body.arg deformed by args/two_body_def.arg, and K(1/2) I
// K(1/2) per slice.  Before each slice SlaterDet2(up, down, one_bodies, NULL, &two_body, N)
// and for n = 0 .. N-1:
//   results/slater2_even.<n+1>         Run(n+1, -1, -1)             (n+1 pairs)
//   results/slater2_odd_<s>.<n+1>      Run(n+1, -1, AccShellCount(s-1)) for every shell s
//   results/slater2_odd_cum_<s>.<n+1>  the same, over the blocks where its real part is > 0
// with the powers log(Re det)^k, k = 1 .. 8, of the accumulated values beside each; one line
// of `%.15e` values per block, per t: corr re, corr im, z1 .. z8 (the odd files: the plain
// corr re, im, then the positive-part corr re, im and z1 .. z8 over unload_period minus the
// number of non-positive values).  results/sources lists "n (opposite-parity index): k1 k2 k3
// log(1 + eps(k))" with the stream's default formatting.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "arg.h"
#include "constants.h"
#include "dispersion.h"
#include "global_job_parameter.h"
#include "lattice.h"
#include "momenta.h"
#include "one_body.h"
#include "propagator.h"
#include "random.h"
#include "slater_det2.h"
#include "sysfunc.h"
#include "sysio.h"
#include "two_body.h"
#include "verbose.h"

namespace {
constexpr int NZ = 8;  // cumulant powers

struct Acc {  // one (n, t) correlator with its log powers
  std::complex<Float> corr = 0.0;
  Float z[NZ] = {0, 0, 0, 0, 0, 0, 0, 0};
  int skipped = 0;
  void add_log(std::complex<Float> v) {
    corr += v;
    const Float l = std::log(v.real());
    Float p = l;
    for (int k = 0; k < NZ; ++k, p *= l) z[k] += p;
  }
};
}  // namespace

int main() {
  const char* fname = "int main()";
  Comms::Initialize();

  VerboseArg verbose_arg("args/verbose.arg");
  DoArg do_arg("args/do.arg");
  LatticeArg lattice_arg("args/lattice.arg");
  EvoArg evo_arg("args/evo.arg");
  RandomArg random_arg("args/random.arg");
  InteractionArg interaction_arg("args/interaction.arg");
  KineticArg kinetic_arg("args/kinetic.arg");
  PropagatorArg prop_arg("args/propagator.arg");
  OneBodyArg one_body_arg("args/one_body.arg");
  TwoBodyArg two_body_arg("args/two_body.arg");
  MomentaArg momenta_arg("args/momenta.arg");
  if (evo_arg.unload_period < 1) ERR.General(fname, "evo_arg.unload_period must be positive.");
  if (one_body_arg.source_type != SOURCE_TYPE_MOM) {
    VRB.Warn(fname, "one_body_arg.source_type not supported:");
    VRB.Warn(fname, "\tusing SOURCE_TYPE_MOM instead.");
    one_body_arg.source_type = SOURCE_TYPE_MOM;
  }
  if (momenta_arg.dispersion_type != kinetic_arg.dispersion_type) {
    VRB.Warn(fname, "momenta_arg.dispersion_type and prop_arg.dispersion_type disagree:");
    VRB.Warn(fname, "\tSetting momenta_arg.dispersion_type to prop_arg.dispersion_type.");
    momenta_arg.dispersion_type = kinetic_arg.dispersion_type;
  }
  if (momenta_arg.mass != kinetic_arg.mass) {
    VRB.Warn(fname, "momenta_arg.mass and prop_arg.mass disagree:");
    VRB.Warn(fname, "\tSetting momenta_arg.mass to prop_arg.mass.");
    momenta_arg.mass = kinetic_arg.mass;
  }

  VRB.SetLevel(verbose_arg);
  GJP.Initialize(do_arg);
  Random rng(random_arg);
  Lattice lattice(lattice_arg, &rng);
  Interaction interaction(&lattice, interaction_arg, 1.0);
  Kinetic kinetic(kinetic_arg, 0.5);
  std::vector<Hamiltonian*> hamiltonians = {&kinetic, &interaction, &kinetic};

  Momenta momenta(momenta_arg);
  const int N = momenta.MomentaCount();
  const int T = GJP.Tsites();
  const int S = momenta.NumShells();
  if (!Comms::Rank()) {  // the stream's default formatting, as main-cum.C:129-143 writes it
    Dispersion dispersion(kinetic_arg.dispersion_type, kinetic_arg.mass, CUTOFF_TYPE_HARD);
    std::ofstream file("results/sources", std::ios_base::app);
    for (int n = 0; n < N; ++n) {
      const std::vector<int> k = momenta.GetMomentum(n);
      file << n << " (" << momenta.OppositeParityIndex(n) << "): " << k[0] << " " << k[1]
           << " " << k[2] << " " << std::log(1 + dispersion.Get(k[0], k[1], k[2])) << "\n";
    }
  }

  std::vector<OneBody*> one_bodies(N);
  for (int n = 0; n < N; ++n) {
    const std::vector<int> k = momenta.GetMomentum(n);
    one_bodies[n] = new OneBody(one_body_arg);
    one_bodies[n]->Set(k[0], k[1], k[2]);
  }
  TwoBody two_body(two_body_arg);
  two_body.Deform("args/two_body_def.arg");

  // files: even per n; odd and odd_cum per (shell, n) for n at or beyond the shell's start
  std::vector<FILE*> efile(N), ofile((size_t)S * N, nullptr), ocfile((size_t)S * N, nullptr);
  char name[99];
  for (int n = 0; n < N; ++n) {
    std::snprintf(name, sizeof(name), "results/slater2_even.%d", n + 1);
    efile[n] = Fopen(IO_TYPE_ROOT, name, "a");
    for (int s = 0; s < S; ++s) {
      if (n < momenta.AccShellCount(s - 1)) continue;
      std::snprintf(name, sizeof(name), "results/slater2_odd_%d.%d", s, n + 1);
      ofile[(size_t)s * N + n] = Fopen(IO_TYPE_ROOT, name, "a");
      std::snprintf(name, sizeof(name), "results/slater2_odd_cum_%d.%d", s, n + 1);
      ocfile[(size_t)s * N + n] = Fopen(IO_TYPE_ROOT, name, "a");
    }
  }

  prop_arg.n_fermions = N;
  Propagator propagator(prop_arg);
  std::vector<Float*> uprops(N), dprops(N);
  for (int n = 0; n < N; ++n) {
    uprops[n] = propagator.Get(n);
    dprops[momenta.OppositeParityIndex(n)] = propagator.Get(n);
  }

  const int U = evo_arg.unload_period;
  std::vector<Acc> even((size_t)N * T), odd((size_t)S * N * T), odd_cum((size_t)S * N * T);
  for (int j = 0; j < evo_arg.configurations; ++j) {
    rng.WriteState(random_arg.file_stem.c_str());
    for (int k = 0; k < U; ++k) {
      for (int n = 0; n < N; ++n) propagator.Set(one_bodies[n]->Get(), n);
      for (int t = 0; t < T; ++t) {
        SlaterDet2 slater_det(uprops.data(), dprops.data(), one_bodies.data(), nullptr,
                              &two_body, N);  // measure first
        for (int n = 0; n < N; ++n) {
          even[(size_t)n * T + t].add_log(slater_det.Run(n + 1, -1, -1));
          for (int s = 0; s < S; ++s) {
            const int t0 = momenta.AccShellCount(s - 1);
            if (n < t0) continue;
            const std::complex<Float> v = slater_det.Run(n + 1, -1, t0);
            const size_t i = ((size_t)s * N + n) * T + t;
            odd[i].corr += v;
            if (v.real() > 0) odd_cum[i].add_log(v);
            else odd_cum[i].skipped += 1;
          }
        }
        for (Hamiltonian* h : hamiltonians) h->Initialize();
        propagator.Run(hamiltonians);
        lattice.Refresh();
      }
    }
    // block averages over the ranks (Comms::GlobalSum / Size), then reset
    const Float ranks = (Float)Comms::Size();
    auto put_acc = [&](FILE* f, Acc& a, Float m) {
      const std::complex<Float> c = Comms::GlobalSum(a.corr / m) / ranks;
      Fprintf(f, "%.*e %.*e ", PREC, c.real(), PREC, c.imag());
      for (int q = 0; q < NZ; ++q)
        Fprintf(f, "%.*e ", PREC, (Comms::GlobalSum(std::complex<Float>(a.z[q] / m, 0.0)) /
                                   ranks).real());
      a = Acc();
    };
    for (int n = 0; n < N; ++n) {
      for (int t = 0; t < T; ++t) put_acc(efile[n], even[(size_t)n * T + t], (Float)U);
      Fprintf(efile[n], "\n");
      for (int s = 0; s < S; ++s) {
        if (n < momenta.AccShellCount(s - 1)) continue;
        FILE* fo = ofile[(size_t)s * N + n];
        FILE* fc = ocfile[(size_t)s * N + n];
        for (int t = 0; t < T; ++t) {
          const size_t i = ((size_t)s * N + n) * T + t;
          const std::complex<Float> c = Comms::GlobalSum(odd[i].corr / (Float)U) / ranks;
          Fprintf(fo, "%.*e %.*e ", PREC, c.real(), PREC, c.imag());
          odd[i] = Acc();
          put_acc(fc, odd_cum[i], (Float)(U - odd_cum[i].skipped));
        }
        Fprintf(fo, "\n");
        Fprintf(fc, "\n");
      }
      if (j % FLUSHRATE == 0) Fflush(efile[n]);
    }
  }
  for (int n = 0; n < N; ++n) delete one_bodies[n];
  for (FILE* f : efile) Fclose(f);
  for (FILE* f : ofile) Fclose(f);
  for (FILE* f : ocfile) Fclose(f);
  Comms::Finalize();
  return EXIT_SUCCESS;
}
