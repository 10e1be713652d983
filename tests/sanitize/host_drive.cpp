// Drives the host-only parts of the class layer (nrcps_amd/host: argument files, the job
// parameters, Dispersion, Momenta, ranlxs and its state files, Gaussian draws, the special
// functions, Vector3) for a build with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_sanitize.py).  It runs in a directory holding args/ and rngs/; nothing here
// touches the device.  Results are not checked (the GPU driver tests pin them).
#include <cstdio>
#include <vector>

#include "arg.h"
#include "dispersion.h"
#include "global_job_parameter.h"
#include "momenta.h"
#include "random.h"
#include "special_functions.h"
#include "vector3.h"
#include "verbose.h"

int main() {
  VerboseArg verbose_arg("args/verbose.arg");
  VRB.SetLevel(verbose_arg);
  DoArg do_arg("args/do.arg");
  GJP.Initialize(do_arg);
  EvoArg evo_arg("args/evo.arg");
  KineticArg kinetic_arg("args/kinetic.arg");
  InteractionArg interaction_arg("args/interaction.arg");
  MomentaArg momenta_arg("args/momenta.arg");
  OneBodyArg one_body_arg("args/one_body.arg");
  double acc = 0.0;
  for (DispersionType t : {DISPERSION_TYPE_STANDARD, DISPERSION_TYPE_QUADRATIC,
                           DISPERSION_TYPE_PERFECT})
    for (CutoffType c : {CUTOFF_TYPE_NONE, CUTOFF_TYPE_HARD}) {
      Dispersion d(t, 1.18343, c);
      for (int i = 0; i < GJP.Xsites(); ++i) acc += d.Get(i, 1, 2);
    }
  Momenta momenta(momenta_arg);
  for (int n = 0; n < momenta.MomentaCount(); ++n) {
    const std::vector<int> k = momenta.GetMomentum(n);
    acc += k[0] + momenta.OppositeParityIndex(n);
  }
  for (int s = 0; s < momenta.NumShells(); ++s) acc += momenta.ShellCount(s) + momenta.AccShellCount(s - 1);
  for (int level = 0; level <= 2; ++level) {
    RandomArg random_arg;
    random_arg.random_type = level == 0 ? RANDOM_TYPE_RANLXS0
                             : level == 1 ? RANDOM_TYPE_RANLXS1 : RANDOM_TYPE_RANLXS2;
    random_arg.seed_type = SEED_TYPE_INPUT;
    random_arg.seed = 7012 + level;
    random_arg.file_stem = "rngs/ranlxs";
    Random rng(random_arg);
    for (int i = 0; i < 5000; ++i) acc += rng.Uniform() + rng.Z(2) + rng.Gauss(1.0);
    rng.WriteState(random_arg.file_stem.c_str());
    Random again(random_arg);
    again.ReadState(random_arg.file_stem.c_str());
    acc += again.Uniform();
  }
  for (int i = -50; i <= 50; ++i) acc += DawsonF(0.2 * i) + HermiteH(i & 7, 0.1 * i);
  Float h[10];
  HermiteH(h, 9, 0.7);
  const Vector3 a(1, 2, 3), b = 2 * a;
  acc += b.GetX() + Vector3ToIndex(IndexToVector3(GJP.Vol() - 1)) + IndexToVector3(5).GetZ();
  std::printf("host clean %g\n", acc);
  return 0;
}
