// slater_det.cpp -- the Slater-determinant correlator driver on the MI355X class layer.
//
// Runs the production slice loop with the reference's class API and file formats
// (production/slater_det/main.C:25-264): arguments from args/*.arg in the working directory,
// sources and sinks from Momenta, K(1/2) I K(1/2) per slice, every leading minor measured
// before each slice, results/slater.<n> holding one line of `%.15e %.15e ` pairs (re, im of
// the accumulated |det|^2 per t, divided by the block size) per block, and the RNG state in
// <file_stem>.<rank> before each block.
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
#include "slater_det.h"
#include "sysfunc.h"
#include "sysio.h"
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
  MomentaArg momenta_arg("args/momenta.arg");
  if (evo_arg.unload_period < 1) ERR.General(fname, "evo_arg.unload_period must be positive.");
  // production/slater_det/main.C:74-95: momentum sources, and the sinks follow the kinetic
  // operator's dispersion and mass
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
  if (!Comms::Rank()) {  // results/sources (main.C:122-140): index (partner): k log(1 + eps)
    Dispersion dispersion(kinetic_arg.dispersion_type, kinetic_arg.mass, CUTOFF_TYPE_HARD);
    std::ofstream file("results/sources", std::ios_base::app);
    for (int n = 0; n < N; ++n) {
      const std::vector<int> k = momenta.GetMomentum(n);
      file << n << " (" << momenta.OppositeParityIndex(n) << "): " << k[0] << " " << k[1]
           << " " << k[2] << " " << std::log(1 + dispersion.Get(k[0], k[1], k[2])) << "\n";
    }
  }
  std::vector<OneBody*> sinks(N);
  for (int n = 0; n < N; ++n) {
    const std::vector<int> k = momenta.GetMomentum(n);
    sinks[n] = new OneBody(one_body_arg);
    sinks[n]->Set(k[0], k[1], k[2]);
  }

  std::vector<FILE*> out(N);
  for (int n = 0; n < N; ++n) {
    char name[64];
    std::snprintf(name, sizeof(name), "results/slater.%d", n + 1);
    out[n] = Fopen(IO_TYPE_ROOT, name, "a");
  }

  prop_arg.n_fermions = N;
  Propagator propagator(prop_arg);
  std::vector<Float*> props(N);
  for (int n = 0; n < N; ++n) props[n] = propagator.Get(n);

  std::vector<std::complex<Float> > corr((size_t)N * T);
  for (int j = 0; j < evo_arg.configurations; ++j) {
    rng.WriteState(random_arg.file_stem.c_str());
    for (int k = 0; k < evo_arg.unload_period; ++k) {
      for (int n = 0; n < N; ++n) propagator.Set(sinks[n]->Get(), n);
      for (int t = 0; t < T; ++t) {
        SlaterDet det(props.data(), sinks.data(), N);  // measure first
        for (int n = 0; n < N; ++n) {
          const std::complex<Float> d = det.Run(n + 1);
          corr[(size_t)n * T + t] += d * std::conj(d);
        }
        for (Hamiltonian* h : hamiltonians) h->Initialize();
        propagator.Run(hamiltonians);
        lattice.Refresh();
      }
    }
    Comms::GlobalSum(corr.data(), N * T);
    const Float block = (Float)Comms::Size() * evo_arg.unload_period;
    for (int n = 0; n < N; ++n) {
      for (int t = 0; t < T; ++t) {
        std::complex<Float>& c = corr[(size_t)n * T + t];
        Fprintf(out[n], "%.*e %.*e ", PREC, c.real() / block, PREC, c.imag() / block);
        c = 0.0;
      }
      Fprintf(out[n], "\n");
      if (j % FLUSHRATE == 0) Fflush(out[n]);
    }
  }
  for (int n = 0; n < N; ++n) {
    delete sinks[n];
    Fclose(out[n]);
  }
  Comms::Finalize();
  return EXIT_SUCCESS;
}
