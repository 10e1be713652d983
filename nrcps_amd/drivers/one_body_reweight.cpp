// one_body_reweight.cpp -- the importance-sampled one-body propagator driver on the MI355X
// class layer.
//
// The loop of production/one_body-reweight/main.C:37-170 with the reference's class API and
// file formats: a single propagator (prop.arg's n_fermions must be one), the source and sink
// read from args/source.arg and args/sink.arg ("k1 k2 k3 re im" lines), slices K(1/2)
// P(1/2) I P(1/2) K(1/2).  Each slice measures <sink|psi(t)>, draws a new field shifted by
// phi0 = 0.0022 (Lattice::Refresh(Float), a Gaussian field moved off zero), evolves, and
// divides the propagator by exp(V phi0 (Re <phi> - phi0 / 2)): the ratio q / p of the
// shifted to the unshifted Gaussian density of the drawn field, so the accumulated overlaps
// are unbiased for the unshifted ensemble.  One line per block of `%.15e %.15e ` pairs per t
// (re, im divided by the block size) into prop.arg's file stem.  The reference's GPU build
// refuses Propagator::Normalize (src/alg/propagator.C:116-133); here it runs on the device.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "arg.h"
#include "constants.h"
#include "global_job_parameter.h"
#include "lattice.h"
#include "one_body.h"
#include "propagator.h"
#include "random.h"
#include "sysfunc.h"
#include "sysio.h"
#include "verbose.h"

int main() {
  const char* fname = "int main()";
  Comms::Initialize();

  VerboseArg verbose_arg("args/verbose.arg");
  VRB.SetLevel(verbose_arg);
  DoArg do_arg("args/do.arg");
  GJP.Initialize(do_arg);
  LatticeArg lattice_arg("args/lattice.arg");
  EvoArg evo_arg("args/evo.arg");
  RandomArg random_arg("args/random.arg");
  PropagatorArg prop_arg("args/prop.arg");
  InteractionArg interaction_arg("args/interaction.arg");
  PotentialArg potential_arg("args/potential.arg");
  KineticArg kinetic_arg("args/kinetic.arg");
  OneBodyArg one_body_arg("args/one_body.arg");
  if (evo_arg.unload_period != 1) ERR.General(fname, "evo_arg.unload_period must be equal to one.");
  if (prop_arg.n_fermions != 1) ERR.General(fname, "Number of prop_arg.n_fermions must equal one.");

  Random rng(random_arg);
  Lattice lattice(lattice_arg, &rng);
  Interaction interaction(&lattice, interaction_arg, 1.0);
  Potential potential(potential_arg, 0.5);
  Kinetic kinetic(kinetic_arg, 0.5);
  std::vector<Hamiltonian*> hamiltonians = {&kinetic, &potential, &interaction, &potential,
                                            &kinetic};

  OneBody source(one_body_arg);
  source.Set("args/source.arg");
  OneBody sink(one_body_arg);
  sink.Set("args/sink.arg");

  FILE* file = Fopen(IO_TYPE_ROOT, prop_arg.file_stem.c_str(), "a");
  const int T = GJP.Tsites();
  std::vector<std::complex<Float> > corr(T, 0.0);
  Propagator prop(prop_arg);
  const Float phi0 = 0.0022;
  for (int j = 0; j < evo_arg.configurations; ++j) {
    rng.WriteState(random_arg.file_stem.c_str());
    for (int k = 0; k < evo_arg.unload_period; ++k) {
      prop.Set(source.Get(), 0);
      for (int t = 0; t < T; ++t) {
        corr[t] += sink.Project(prop.Get(0));
        lattice.Refresh(phi0);
        for (Hamiltonian* h : hamiltonians) h->Initialize();
        prop.Run(hamiltonians);
        prop.Normalize(std::exp(GJP.Vol() * phi0 * (std::real(lattice.Mean()) - 0.5 * phi0)));
      }
    }
    Comms::GlobalSum(corr.data(), T);
    const Float block = (Float)Comms::Size() * evo_arg.unload_period;
    for (int t = 0; t < T; ++t) {
      Fprintf(file, "%.*e %.*e ", PREC, corr[t].real() / block, PREC, corr[t].imag() / block);
      corr[t] = 0.0;
    }
    Fprintf(file, "\n");
    if (j % FLUSHRATE == 0) Fflush(file);
  }
  Fclose(file);
  Comms::Finalize();
  return EXIT_SUCCESS;
}
