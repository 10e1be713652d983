// slater_det3.cpp -- the pair-sink correlator driver on the MI355X class layer.
//
// The measurement loop of production/slater_det3_trapped/main-cum.C:140-295 with the
// reference's class API and file formats: arguments from args/*.arg (two_body.arg for the
// pair wave function), N = propagator.arg's n_fermions SHO sources and one-body sinks whose
// coefficients are read from sources/source.<n> ("n1 n2 n3 re im" lines, as main-cum.C:75-78
// and 139-143 set them), K(1/2) I K(1/2) per slice, or K(1/2) P(1/2) I P(1/2) K(1/2) when
// args/potential.arg is present (main-cum.C:97-106); before each slice SlaterDet3::Initialize(&two_body) and
// for n = 0 .. N-1 the determinants Run(n+1, -1, -1) (n+1 pairs, into results/slater.<2n+2>)
// and Run(n+1, -1, n) (n pairs and one particle, into results/slater.<2n+1>), with the
// powers log|det|^(k+1), k < 6, accumulated into results/z<k+1>.<m>.  One line of
// `%.15e %.15e ` pairs (re, im per t divided by the block size) per block; the cumulant files
// hold the real parts.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "arg.h"
#include "constants.h"
#include "dispersion.h"
#include "global_job_parameter.h"
#include "lattice.h"
#include "one_body.h"
#include "propagator.h"
#include "random.h"
#include "slater_det3.h"
#include "sysfunc.h"
#include "sysio.h"
#include "two_body.h"
#include "verbose.h"

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
  if (evo_arg.unload_period < 1) ERR.General(fname, "evo_arg.unload_period must be positive.");
  if (one_body_arg.source_type != SOURCE_TYPE_SHO) {  // main-cum.C:75-78
    VRB.Warn(fname, "Setting one_body_arg.source_type to SOURCE_TYPE_SHO.");
    one_body_arg.source_type = SOURCE_TYPE_SHO;
  }
  const int N = prop_arg.n_fermions;
  if (N < 1 || N > 35) ERR.General(fname, "propagator_arg.n_fermions must be in 1 .. 35.");

  VRB.SetLevel(verbose_arg);
  GJP.Initialize(do_arg);
  Random rng(random_arg);
  Lattice lattice(lattice_arg, &rng);
  Interaction interaction(&lattice, interaction_arg, 1.0);
  Kinetic kinetic(kinetic_arg, 0.5);
  std::vector<Hamiltonian*> hamiltonians = {&kinetic, &interaction, &kinetic};
  // with args/potential.arg the trapped improved order K, P(1/2), I, P(1/2), K
  // (production/slater_det3_trapped/main-cum.C:97-106)
  Potential* potential = nullptr;
  if (FILE* pf = std::fopen("args/potential.arg", "r")) {
    std::fclose(pf);
    PotentialArg potential_arg("args/potential.arg");
    potential = new Potential(potential_arg, 0.5);
    hamiltonians = {&kinetic, potential, &interaction, potential, &kinetic};
  }

  const int T = GJP.Tsites();
  std::vector<OneBody*> one_bodies(N);
  for (int n = 0; n < N; ++n) {
    char name[64];
    std::snprintf(name, sizeof(name), "sources/source.%d", n);
    one_bodies[n] = new OneBody(one_body_arg);
    one_bodies[n]->Set(name);
  }
  TwoBody two_body(two_body_arg);

  const int Nc = 6;
  std::vector<FILE*> file(2 * N), zfile((size_t)Nc * 2 * N);
  for (int m = 0; m < 2 * N; ++m) {
    char name[64];
    std::snprintf(name, sizeof(name), "results/slater.%d", m + 1);
    file[m] = Fopen(IO_TYPE_ROOT, name, "a");
    for (int nc = 0; nc < Nc; ++nc) {
      std::snprintf(name, sizeof(name), "results/z%d.%d", nc + 1, m + 1);
      zfile[(size_t)nc * 2 * N + m] = Fopen(IO_TYPE_ROOT, name, "a");
    }
  }

  Propagator propagator(prop_arg);
  std::vector<Float*> props(N);
  for (int n = 0; n < N; ++n) props[n] = propagator.Get(n);
  SlaterDet3 slater_det(props.data(), props.data(), one_bodies.data(), nullptr, N);

  std::vector<std::complex<Float> > corr((size_t)2 * N * T), z((size_t)Nc * 2 * N * T);
  auto accumulate = [&](int m, int t, std::complex<Float> slat) {
    corr[(size_t)m * T + t] += slat;
    Float cum = std::log(std::abs(slat)), p = cum;
    for (int nc = 0; nc < Nc; ++nc, p *= cum) z[((size_t)nc * 2 * N + m) * T + t] += p;
  };
  for (int j = 0; j < evo_arg.configurations; ++j) {
    rng.WriteState(random_arg.file_stem.c_str());
    for (int k = 0; k < evo_arg.unload_period; ++k) {
      for (int n = 0; n < N; ++n) propagator.Set(one_bodies[n]->Get(), n);
      for (int t = 0; t < T; ++t) {
        slater_det.Initialize(&two_body);  // measure first
        for (int n = 0; n < N; ++n) {
          accumulate(2 * n + 1, t, slater_det.Run(n + 1, -1, -1));
          accumulate(2 * n, t, slater_det.Run(n + 1, -1, n));
        }
        for (Hamiltonian* h : hamiltonians) h->Initialize();
        propagator.Run(hamiltonians);
        lattice.Refresh();
      }
    }
    Comms::GlobalSum(corr.data(), 2 * N * T);
    Comms::GlobalSum(z.data(), Nc * 2 * N * T);
    const Float block = (Float)Comms::Size() * evo_arg.unload_period;
    for (int m = 0; m < 2 * N; ++m) {
      for (int t = 0; t < T; ++t) {
        std::complex<Float>& c = corr[(size_t)m * T + t];
        Fprintf(file[m], "%.*e %.*e ", PREC, c.real() / block, PREC, c.imag() / block);
        c = 0.0;
        for (int nc = 0; nc < Nc; ++nc) {
          std::complex<Float>& q = z[((size_t)nc * 2 * N + m) * T + t];
          Fprintf(zfile[(size_t)nc * 2 * N + m], "%.*e ", PREC, q.real() / block);
          q = 0.0;
        }
      }
      Fprintf(file[m], "\n");
      for (int nc = 0; nc < Nc; ++nc) Fprintf(zfile[(size_t)nc * 2 * N + m], "\n");
      if (j % FLUSHRATE == 0) {
        Fflush(file[m]);
        for (int nc = 0; nc < Nc; ++nc) Fflush(zfile[(size_t)nc * 2 * N + m]);
      }
    }
  }
  for (int n = 0; n < N; ++n) delete one_bodies[n];
  delete potential;
  for (FILE* f : file) Fclose(f);
  for (FILE* f : zfile) Fclose(f);
  Comms::Finalize();
  return EXIT_SUCCESS;
}
