// util.cpp -- argument files, global job parameters, logging, fatal errors, the single
// process Comms layer and rank-aware stdio of the host class layer.
//
// Reference behaviour: src/utils/arg.C:8-831 (positional whitespace-separated .arg files,
// enums spelled by name, InteractionArg reading couplings until end of file),
// global_job_parameter.C:30-112, verbose.C:10-103, error.C:9-58, sysfunc.C:21-151 (one
// rank here), sysio.C:12-64, vector3.C:151-233.
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iostream>

#include "../../include/nrcps/nrcps_host.h"

GlobalJobParameter GJP;
Verbose VRB;
Error ERR;

// ------------------------------------------------------------------ device context
nrc_context* nrc_host_context() {
  static nrc_context* ctx = nullptr;
  if (!ctx) {
    const char* dev = std::getenv("NRC_DEVICE");
    const int rc = nrc_init(dev ? std::atoi(dev) : 0, &ctx);
    if (rc != NRC_OK) ERR.General("nrc_host_context()", "nrc_init: %s", nrc_error_string(rc));
  }
  return ctx;
}

// ------------------------------------------------------------------ Error
static void fatal_exit() {
  std::fflush(stdout);
  std::exit(EXIT_FAILURE);
}

void Error::FileR(const char* fname, const char* file) {
  std::printf("ERROR in %s: cannot open %s for reading.\n", fname, file);
  fatal_exit();
}
void Error::FileW(const char* fname, const char* file) {
  std::printf("ERROR in %s: cannot open %s for writing.\n", fname, file);
  fatal_exit();
}
void Error::FileA(const char* fname, const char* file) {
  std::printf("ERROR in %s: cannot open %s for appending.\n", fname, file);
  fatal_exit();
}
void Error::General(const char* fname, const char* fmt, ...) {
  std::printf("ERROR in %s: ", fname);
  va_list ap;
  va_start(ap, fmt);
  std::vprintf(fmt, ap);
  va_end(ap);
  std::printf("\n");
  fatal_exit();
}
void Error::NotImplemented(const char* fname, const char* fmt, ...) {
  std::printf("ERROR in %s: not implemented: ", fname);
  va_list ap;
  va_start(ap, fmt);
  std::vprintf(fmt, ap);
  va_end(ap);
  std::printf("\n");
  fatal_exit();
}
void Error::Check(const char* fname, int status) {
  if (status != NRC_OK) General(fname, "%s (status %d)", nrc_error_string(status), status);
}

// ------------------------------------------------------------------ Verbose
Verbose::Verbose()
    : func_level(false), warn_level(true), result_level(true), flow_level(false),
      debug_level(false) {}

void Verbose::SetLevel(const VerboseArg& a) {
  func_level = a.func_level;
  warn_level = a.warn_level;
  result_level = a.result_level;
  flow_level = a.flow_level;
  debug_level = a.debug_level;
}

static void vrb_print(const char* tag, const char* fname, const char* fmt, va_list ap) {
  if (Comms::Rank() != 0) return;
  std::printf("%s %s: ", tag, fname);
  std::vprintf(fmt, ap);
  std::printf("\n");
}

void Verbose::Func(const char* fname) {
  if (func_level && Comms::Rank() == 0) std::printf("FUNC %s\n", fname);
}
#define NRC_VRB_LEVEL(NAME, FLAG, TAG)                   \
  void Verbose::NAME(const char* fname, const char* fmt, ...) { \
    if (!FLAG) return;                                   \
    va_list ap;                                          \
    va_start(ap, fmt);                                   \
    vrb_print(TAG, fname, fmt, ap);                      \
    va_end(ap);                                          \
  }
NRC_VRB_LEVEL(Warn, warn_level, "WARN")
NRC_VRB_LEVEL(Result, result_level, "RESULT")
NRC_VRB_LEVEL(Flow, flow_level, "FLOW")
NRC_VRB_LEVEL(Debug, debug_level, "DEBUG")

void Verbose::LibInfo(const char* fname) {
  if (flow_level && Comms::Rank() == 0) std::printf("FLOW %s: library %s\n", fname, VERSION_STR);
}

// ------------------------------------------------------------------ Comms (one rank)
namespace Comms {
void Initialize() {}
void Finalize() {}
void RaiseError() { fatal_exit(); }
int Rank() { return 0; }
int Size() { return 1; }
void Sync() {}
std::complex<Float> GlobalSum(std::complex<Float> x) { return x; }
void GlobalSum(std::complex<Float>*, int) {}
}  // namespace Comms

// ------------------------------------------------------------------ stdio
FILE* Fopen(IoType type, const char* filename, const char* mode) {
  if (type == IO_TYPE_ALL) {
    std::string f = std::string(filename) + "." + std::to_string(Comms::Rank());
    FILE* fp = std::fopen(f.c_str(), mode);
    if (!fp) ERR.FileR("Fopen", f.c_str());
    return fp;
  }
  FILE* fp = std::fopen(filename, mode);
  if (!fp) ERR.FileW("Fopen", filename);
  return fp;
}
int Fclose(FILE* s) { return s ? std::fclose(s) : 0; }
int Fprintf(FILE* s, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  const int r = std::vfprintf(s, fmt, ap);
  va_end(ap);
  return r;
}
size_t Fwrite(const void* p, size_t size, size_t n, FILE* s) { return std::fwrite(p, size, n, s); }
size_t Fread(void* p, size_t size, size_t n, FILE* s) { return std::fread(p, size, n, s); }
int Fflush(FILE* s) { return std::fflush(s); }

// ------------------------------------------------------------------ GlobalJobParameter
GlobalJobParameter::GlobalJobParameter()
    : t_sites(4), x_sites(4), y_sites(4), z_sites(4), x_bc(BOUNDARY_TYPE_PRD),
      y_bc(BOUNDARY_TYPE_PRD), z_bc(BOUNDARY_TYPE_PRD), cutoff(PI) {}

void GlobalJobParameter::Initialize(const DoArg& d) {
  const char* fname = "void GlobalJobParameter::Initialize(const DoArg &)";
  t_sites = d.t_sites;
  x_sites = d.x_sites;
  y_sites = d.y_sites;
  z_sites = d.z_sites;
  if (x_sites % 2) ERR.General(fname, "x_sites not divisible by two.");
  if (y_sites % 2) ERR.General(fname, "y_sites not divisible by two.");
  if (z_sites % 2) ERR.General(fname, "z_sites not divisible by two.");
  x_bc = d.x_boundary_type;
  y_bc = d.y_boundary_type;
  z_bc = d.z_boundary_type;
  if (d.cutoff <= 0.0) ERR.General(fname, "Cutoff factor is not positive.");
  cutoff = d.cutoff * PI;
}

bool GlobalJobParameter::APBCQuery() const {
  return x_bc == BOUNDARY_TYPE_APRD || y_bc == BOUNDARY_TYPE_APRD || z_bc == BOUNDARY_TYPE_APRD;
}

void GlobalJobParameter::GetBcShift(Float* bc) const {
  bc[0] = x_bc == BOUNDARY_TYPE_APRD ? 0.5 : 0.0;
  bc[1] = y_bc == BOUNDARY_TYPE_APRD ? 0.5 : 0.0;
  bc[2] = z_bc == BOUNDARY_TYPE_APRD ? 0.5 : 0.0;
}

// ------------------------------------------------------------------ Vector3 helpers
const Vector3 operator*(int a, const Vector3& v) {
  return Vector3(a * v.GetX(), a * v.GetY(), a * v.GetZ());
}

static int wrap(int a, int n) { return ((a % n) + n) % n; }

int Vector3ToIndex(const Vector3& v) {
  const int X = GJP.Xsites(), Y = GJP.Ysites(), Z = GJP.Zsites();
  return wrap(v.GetZ(), Z) + Z * (wrap(v.GetY(), Y) + Y * wrap(v.GetX(), X));
}

const Vector3 IndexToVector3(int i) {
  if (i < 0 || i >= GJP.Vol()) ERR.General("IndexToVector3(int)", "Index out of bounds.");
  const int Y = GJP.Ysites(), Z = GJP.Zsites();
  return Vector3(i / (Y * Z), (i / Z) % Y, i % Z);
}

// ------------------------------------------------------------------ Arg files
void Arg::OpenFile(std::string f) {
  file.open(f.c_str());
  if (!file) ERR.FileR("void Arg::OpenFile(std::string)", f.c_str());
}
void Arg::CloseFile() { file.close(); }

std::string Arg::token(const char* what) {
  std::string s;
  file >> s;
  (void)what;
  return s;
}

#define NRC_READ_NUM(T)                   \
  int Arg::ReadLine(T& p) {               \
    file >> p;                            \
    return file.eof() ? 1 : 0;            \
  }
NRC_READ_NUM(int)
NRC_READ_NUM(long)
NRC_READ_NUM(Float)
NRC_READ_NUM(std::string)

int Arg::ReadLine(bool& p) {
  const std::string s = token("bool");
  if (s == "1") p = true;
  else if (s == "0") p = false;
  else ERR.General("int Arg::ReadLine(bool &)", "Unable to recognize bool");
  return file.eof() ? 1 : 0;
}

// enums are spelled by name in the files (arg.C:58-335)
template <typename E>
static bool enum_from(const std::string& s, E& out, const char* const* names, int n) {
  for (int i = 0; i < n; ++i)
    if (s == names[i]) {
      out = (E)i;
      return true;
    }
  return false;
}

#define NRC_READ_ENUM(E, ...)                                                        \
  int Arg::ReadLine(E& p) {                                                          \
    static const char* const names[] = {__VA_ARGS__};                                \
    const std::string s = token(#E);                                                 \
    if (!enum_from(s, p, names, (int)(sizeof(names) / sizeof(names[0]))))            \
      ERR.General("int Arg::ReadLine(" #E " &)", "Unable to recognize %s", s.c_str()); \
    return file.eof() ? 1 : 0;                                                       \
  }
NRC_READ_ENUM(RandomType, "RANDOM_TYPE_RAN0", "RANDOM_TYPE_RAN1", "RANDOM_TYPE_RAN2",
              "RANDOM_TYPE_RAN3", "RANDOM_TYPE_RANLXS0", "RANDOM_TYPE_RANLXS1",
              "RANDOM_TYPE_RANLXS2", "RANDOM_TYPE_RANLXD1", "RANDOM_TYPE_RANLXD2")
NRC_READ_ENUM(SeedType, "SEED_TYPE_DEFAULT", "SEED_TYPE_INPUT", "SEED_TYPE_TIME",
              "SEED_TYPE_FILE")
NRC_READ_ENUM(FieldType, "FIELD_TYPE_GAUSS", "FIELD_TYPE_COMPLEXGAUSS", "FIELD_TYPE_ZTWO",
              "FIELD_TYPE_ZTHREE")
NRC_READ_ENUM(SourceType, "SOURCE_TYPE_MOM", "SOURCE_TYPE_SHO")
NRC_READ_ENUM(BoundaryType, "BOUNDARY_TYPE_PRD", "BOUNDARY_TYPE_APRD")
NRC_READ_ENUM(DispersionType, "DISPERSION_TYPE_STANDARD", "DISPERSION_TYPE_QUADRATIC",
              "DISPERSION_TYPE_PERFECT")
NRC_READ_ENUM(CutoffType, "CUTOFF_TYPE_NONE", "CUTOFF_TYPE_HARD")
NRC_READ_ENUM(PotentialType, "POTENTIAL_TYPE_LIN", "POTENTIAL_TYPE_EXP")
NRC_READ_ENUM(PotentialForm, "POTENTIAL_FORM_NONE", "POTENTIAL_FORM_HARMONIC",
              "POTENTIAL_FORM_COULOMB")
NRC_READ_ENUM(DetType, "DET_TYPE_STANDARD", "DET_TYPE_LOG")
NRC_READ_ENUM(InteractionType, "INTERACTION_TYPE_NONE", "INTERACTION_TYPE_XISQMINUSONE",
              "INTERACTION_TYPE_ONEMINUSXIINVSQ")
NRC_READ_ENUM(IoType, "IO_TYPE_ROOT", "IO_TYPE_ALL")
NRC_READ_ENUM(WavefuncType, "WAVEFUNC_TYPE_NONE", "WAVEFUNC_TYPE_UNIFORM", "WAVEFUNC_TYPE_GND",
              "WAVEFUNC_TYPE_TGND", "WAVEFUNC_TYPE_PAIR1", "WAVEFUNC_TYPE_PAIR2")

DoArg::DoArg()
    : t_sites(4), x_sites(4), y_sites(4), z_sites(4), x_boundary_type(BOUNDARY_TYPE_PRD),
      y_boundary_type(BOUNDARY_TYPE_PRD), z_boundary_type(BOUNDARY_TYPE_PRD), cutoff(1.0) {}
void DoArg::Decode(std::string f) {
  OpenFile(f);
  ReadLine(t_sites);
  ReadLine(x_sites);
  ReadLine(y_sites);
  ReadLine(z_sites);
  ReadLine(x_boundary_type);
  ReadLine(y_boundary_type);
  ReadLine(z_boundary_type);
  ReadLine(cutoff);
  CloseFile();
}

LatticeArg::LatticeArg() : field_type(FIELD_TYPE_GAUSS) {}
void LatticeArg::Decode(std::string f) {
  OpenFile(f);
  ReadLine(field_type);
  CloseFile();
}

RandomArg::RandomArg()
    : random_type(RANDOM_TYPE_RANLXS2), seed_type(SEED_TYPE_DEFAULT), seed(1),
      file_stem("rng") {}
void RandomArg::Decode(std::string f) {
  OpenFile(f);
  ReadLine(random_type);
  ReadLine(seed_type);
  ReadLine(seed);
  ReadLine(file_stem);
  CloseFile();
}

PropagatorArg::PropagatorArg() : n_fermions(1), file_stem("prop") {}
void PropagatorArg::Decode(std::string f) {
  OpenFile(f);
  ReadLine(n_fermions);
  ReadLine(file_stem);
  CloseFile();
}

OneBodyArg::OneBodyArg() : source_type(SOURCE_TYPE_MOM), lambda1(1.0), lambda2(1.0), lambda3(1.0) {}
void OneBodyArg::Decode(std::string f) {
  OpenFile(f);
  ReadLine(source_type);
  ReadLine(lambda1);
  ReadLine(lambda2);
  ReadLine(lambda3);
  CloseFile();
}

VerboseArg::VerboseArg()
    : func_level(false), warn_level(true), result_level(true), flow_level(false),
      debug_level(false) {}
void VerboseArg::Decode(std::string f) {
  OpenFile(f);
  ReadLine(func_level);
  ReadLine(warn_level);
  ReadLine(result_level);
  ReadLine(flow_level);
  ReadLine(debug_level);
  CloseFile();
}

EvoArg::EvoArg() : start(0), unload_period(1), configurations(1) {}
void EvoArg::Decode(std::string f) {
  OpenFile(f);
  ReadLine(start);
  ReadLine(unload_period);
  ReadLine(configurations);
  CloseFile();
}

MomentaArg::MomentaArg() : dispersion_type(DISPERSION_TYPE_PERFECT), fermi_energy(1.0), mass(1.0) {}
void MomentaArg::Decode(std::string f) {
  OpenFile(f);
  ReadLine(dispersion_type);
  ReadLine(fermi_energy);
  ReadLine(mass);
  CloseFile();
}

InteractionArg::InteractionArg()
    : mass(1.0), interaction_type(INTERACTION_TYPE_NONE), num_couplings(0) {}
void InteractionArg::Decode(std::string f) {
  OpenFile(f);
  ReadLine(mass);
  ReadLine(interaction_type);
  couplings.clear();
  Float c;
  while (true) {  // couplings until end of file (arg.C:743-747)
    if (ReadLine(c)) break;
    couplings.push_back(c);
  }
  CloseFile();
  num_couplings = (int)couplings.size();
}

PotentialArg::PotentialArg()
    : potential_form(POTENTIAL_FORM_NONE), potential_type(POTENTIAL_TYPE_LIN),
      spring_constant1(0.0), spring_constant2(0.0), spring_constant3(0.0) {}
void PotentialArg::Decode(std::string f) {
  OpenFile(f);
  ReadLine(potential_form);
  ReadLine(potential_type);
  ReadLine(spring_constant1);
  ReadLine(spring_constant2);
  ReadLine(spring_constant3);
  CloseFile();
}

// arg.C:523-558
TwoBodyArg::TwoBodyArg()
    : wavefunc_type(WAVEFUNC_TYPE_UNIFORM), dispersion_type1(DISPERSION_TYPE_STANDARD),
      dispersion_type2(DISPERSION_TYPE_STANDARD), mass1(1.0), mass2(1.0), lambda(1.0),
      file_stem("two_body") {}
void TwoBodyArg::Decode(std::string f) {
  OpenFile(f);
  ReadLine(wavefunc_type);
  ReadLine(dispersion_type1);
  ReadLine(dispersion_type2);
  ReadLine(mass1);
  ReadLine(mass2);
  ReadLine(lambda);
  ReadLine(file_stem);
  CloseFile();
}

KineticArg::KineticArg() : dispersion_type(DISPERSION_TYPE_STANDARD), mass(1.0) {}
void KineticArg::Decode(std::string f) {
  OpenFile(f);
  ReadLine(dispersion_type);
  ReadLine(mass);
  CloseFile();
}
