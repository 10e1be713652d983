// nrcps_host.h -- the host class layer of nrCPS's propagator / Slater-determinant path on
// MI355X, with the reference's class names, signatures and semantics so that its production
// drivers compile against it unchanged.  Device work goes through the C ABI of nrcps.h.
//
// Reference interfaces (file:line in the nrCPS tree):
//   config.h:37-48 (Float, PREC), enum.h:6-133, constants.h:9-14, arg.h:11-213,
//   global_job_parameter.h:11-41, vector3.h:62-66, verbose.h:10-38, error.h:9-18,
//   sysfunc.h:8-25, sysio.h:8-13, fourier.h:7-45, dispersion.h:5-18, hamiltonian.h:9-93,
//   propagator.h:10-28, lattice.h:8-35, random.h:6-26, uniform_deviates.h:163-200,
//   one_body.h:9-41, slater_det.h:9-27, ludcmp.h:5, momenta.h:8-29.
//
// Pointers returned by Fourier::Get, Propagator::Get, Lattice::Get and OneBody::Get are
// device pointers, as in the reference's GPU build; Project / SlaterDet take device pointers.
#ifndef NRCPS_HOST_H
#define NRCPS_HOST_H

#include <complex>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../nrcps.h"

// ------------------------------------------------------------------ precision (config.h)
#define VERSION_STR "nrCPS_v2_7_6 / MI355X"
#ifdef NRC_USE_SINGLE
typedef float Float;
const int PREC = 6;
#define NRC_FLOAT_PREC NRC_PREC_F32
#else
typedef double Float;
const int PREC = 15;
#define NRC_FLOAT_PREC NRC_PREC_F64
#endif

// device complex of the build precision (fourier_typedefs.h: cufftComplex / DoubleComplex)
struct fftComplex {
  Float x, y;
};

// ------------------------------------------------------------------ enums (enum.h)
enum RandomType { RANDOM_TYPE_RAN0, RANDOM_TYPE_RAN1, RANDOM_TYPE_RAN2, RANDOM_TYPE_RAN3,
                  RANDOM_TYPE_RANLXS0, RANDOM_TYPE_RANLXS1, RANDOM_TYPE_RANLXS2,
                  RANDOM_TYPE_RANLXD1, RANDOM_TYPE_RANLXD2 };
enum SeedType { SEED_TYPE_DEFAULT, SEED_TYPE_INPUT, SEED_TYPE_TIME, SEED_TYPE_FILE };
enum FieldType { FIELD_TYPE_GAUSS, FIELD_TYPE_COMPLEXGAUSS, FIELD_TYPE_ZTWO, FIELD_TYPE_ZTHREE };
enum SourceType { SOURCE_TYPE_MOM, SOURCE_TYPE_SHO };
enum WavefuncType { WAVEFUNC_TYPE_NONE, WAVEFUNC_TYPE_UNIFORM, WAVEFUNC_TYPE_GND,
                    WAVEFUNC_TYPE_TGND, WAVEFUNC_TYPE_PAIR1, WAVEFUNC_TYPE_PAIR2 };
enum BoundaryType { BOUNDARY_TYPE_PRD = 0, BOUNDARY_TYPE_APRD = 1 };
enum DispersionType { DISPERSION_TYPE_STANDARD, DISPERSION_TYPE_QUADRATIC,
                      DISPERSION_TYPE_PERFECT };
enum CutoffType { CUTOFF_TYPE_NONE, CUTOFF_TYPE_HARD };
enum PotentialType { POTENTIAL_TYPE_LIN, POTENTIAL_TYPE_EXP };
enum PotentialForm { POTENTIAL_FORM_NONE, POTENTIAL_FORM_HARMONIC, POTENTIAL_FORM_COULOMB };
enum DetType { DET_TYPE_STANDARD, DET_TYPE_LOG };
enum InteractionType { INTERACTION_TYPE_NONE, INTERACTION_TYPE_XISQMINUSONE,
                       INTERACTION_TYPE_ONEMINUSXIINVSQ };
enum IoType { IO_TYPE_ROOT, IO_TYPE_ALL };

// ------------------------------------------------------------------ constants (constants.h)
const int FLUSHRATE = 10;
const Float PI = 3.1415926535897932;
const Float TWOPI = 6.2831853071795864;
const Float PIM4 = 0.7511255444649425;  // pi^(-1/4)
const Float PISQ = 9.8696044010893586;

// ------------------------------------------------------------------ Vector3 (vector3.h)
class Vector3 {
 public:
  explicit Vector3(int x, int y, int z) : x_(x), y_(y), z_(z) {}
  explicit Vector3(int a = 0) : x_(a), y_(a), z_(a) {}
  int GetX() const { return x_; }
  int GetY() const { return y_; }
  int GetZ() const { return z_; }
  void Set(int x, int y, int z) { x_ = x; y_ = y; z_ = z; }

 private:
  int x_, y_, z_;
};
const Vector3 operator*(int, const Vector3&);
int Vector3ToIndex(const Vector3&);
const Vector3 IndexToVector3(int);

// ------------------------------------------------------------------ args (arg.h)
class Arg {
 public:
  Arg() {}
  virtual ~Arg() {}
  virtual void Decode(std::string) = 0;
  void OpenFile(std::string);
  void CloseFile();
  // whitespace-separated tokens; returns 1 at end of file
  int ReadLine(int&);
  int ReadLine(long&);
  int ReadLine(Float&);
  int ReadLine(bool&);
  int ReadLine(std::string&);
  int ReadLine(RandomType&);
  int ReadLine(SeedType&);
  int ReadLine(FieldType&);
  int ReadLine(SourceType&);
  int ReadLine(BoundaryType&);
  int ReadLine(DispersionType&);
  int ReadLine(CutoffType&);
  int ReadLine(PotentialType&);
  int ReadLine(PotentialForm&);
  int ReadLine(DetType&);
  int ReadLine(InteractionType&);
  int ReadLine(IoType&);
  int ReadLine(WavefuncType&);

 private:
  std::ifstream file;
  std::string token(const char* what);
};

#define NRC_ARG_CTORS(C)                     \
  explicit C();                              \
  explicit C(std::string f) : C() { Decode(f); } \
  void Decode(std::string);

class DoArg : public Arg {
 public:
  NRC_ARG_CTORS(DoArg)
  int t_sites, x_sites, y_sites, z_sites;
  BoundaryType x_boundary_type, y_boundary_type, z_boundary_type;
  Float cutoff;
};
class LatticeArg : public Arg {
 public:
  NRC_ARG_CTORS(LatticeArg)
  FieldType field_type;
};
class RandomArg : public Arg {
 public:
  NRC_ARG_CTORS(RandomArg)
  RandomType random_type;
  SeedType seed_type;
  long seed;
  std::string file_stem;
};
class PropagatorArg : public Arg {
 public:
  NRC_ARG_CTORS(PropagatorArg)
  int n_fermions;
  std::string file_stem;
};
class OneBodyArg : public Arg {
 public:
  NRC_ARG_CTORS(OneBodyArg)
  SourceType source_type;
  Float lambda1, lambda2, lambda3;
};
class VerboseArg : public Arg {
 public:
  NRC_ARG_CTORS(VerboseArg)
  bool func_level, warn_level, result_level, flow_level, debug_level;
};
class EvoArg : public Arg {
 public:
  NRC_ARG_CTORS(EvoArg)
  int start, unload_period, configurations;
};
class MomentaArg : public Arg {
 public:
  NRC_ARG_CTORS(MomentaArg)
  DispersionType dispersion_type;
  Float fermi_energy, mass;
};
class InteractionArg : public Arg {
 public:
  NRC_ARG_CTORS(InteractionArg)
  Float mass;
  InteractionType interaction_type;
  std::vector<Float> couplings;
  int num_couplings;
};
class PotentialArg : public Arg {
 public:
  NRC_ARG_CTORS(PotentialArg)
  PotentialForm potential_form;
  PotentialType potential_type;
  Float spring_constant1, spring_constant2, spring_constant3;
};
class KineticArg : public Arg {
 public:
  NRC_ARG_CTORS(KineticArg)
  DispersionType dispersion_type;
  Float mass;
};
class TwoBodyArg : public Arg {
 public:
  NRC_ARG_CTORS(TwoBodyArg)
  WavefuncType wavefunc_type;
  DispersionType dispersion_type1, dispersion_type2;
  Float mass1, mass2, lambda;
  std::string file_stem;
};

// ------------------------------------------------------------------ globals
class GlobalJobParameter {
 public:
  GlobalJobParameter();
  void Initialize(const DoArg&);
  int Tsites() const { return t_sites; }
  int Xsites() const { return x_sites; }
  int Ysites() const { return y_sites; }
  int Zsites() const { return z_sites; }
  int Vol() const { return x_sites * y_sites * z_sites; }
  Vector3 GetDimensions() const { return Vector3(x_sites, y_sites, z_sites); }
  BoundaryType Xbc() const { return x_bc; }
  BoundaryType Ybc() const { return y_bc; }
  BoundaryType Zbc() const { return z_bc; }
  bool APBCQuery() const;
  void GetBcShift(Float*) const;
  Float Cutoff() const { return cutoff; }
  Float CutoffSq() const { return cutoff * cutoff; }

 private:
  int t_sites, x_sites, y_sites, z_sites;
  BoundaryType x_bc, y_bc, z_bc;
  Float cutoff;
};
extern GlobalJobParameter GJP;

class Verbose {
 public:
  Verbose();
  void SetLevel(const VerboseArg&);
  void Func(const char*);
  void Warn(const char*, const char*, ...);
  void Result(const char*, const char*, ...);
  void Flow(const char*, const char*, ...);
  void Debug(const char*, const char*, ...);
  void LibInfo(const char*);

 private:
  bool func_level, warn_level, result_level, flow_level, debug_level;
};
extern Verbose VRB;

class Error {
 public:
  void FileR(const char*, const char*);
  void FileW(const char*, const char*);
  void FileA(const char*, const char*);
  void General(const char*, const char*, ...);
  void NotImplemented(const char*, const char*, ...);
  // a non-zero C-ABI status terminates like the reference's fatal errors
  void Check(const char* fname, int status);
};
extern Error ERR;

// one process, one GPU (the reference's MPI layer reduces over ranks: sysfunc.C:21-151)
namespace Comms {
void Initialize();
void Finalize();
void RaiseError();
int Rank();
int Size();
void Sync();
std::complex<Float> GlobalSum(std::complex<Float>);
void GlobalSum(std::complex<Float>*, int);
}  // namespace Comms

FILE* Fopen(IoType type, const char* filename, const char* mode);
int Fclose(FILE* stream);
int Fprintf(FILE* stream, const char* format, ...);
size_t Fwrite(const void* ptr, size_t size, size_t n, FILE* stream);
size_t Fread(void* ptr, size_t size, size_t n, FILE* stream);
int Fflush(FILE* stream);

// the device context every class below uses (created on first use, device 0 or
// $NRC_DEVICE)
nrc_context* nrc_host_context();

// ------------------------------------------------------------------ random numbers
class UniformDeviate {
 public:
  virtual ~UniformDeviate() {}
  virtual void Init(long) = 0;
  virtual Float Run() = 0;
  virtual int StateSize() = 0;
  virtual void GetState(long*) = 0;
  virtual void SetState(long*) = 0;
  virtual RandomType TypeQ() = 0;
};

// Luscher's ranlxs: 24-bit subtract-with-borrow with luxury decimation (ranlxs.C:49-265)
class Ranlxs : public UniformDeviate {
 public:
  explicit Ranlxs(int level);
  void Init(long);
  Float Run();
  int StateSize() { return 105; }
  void GetState(long*);
  void SetState(long*);
  RandomType TypeQ();

 private:
  void Update();
  int num[96];  // pair k, half h, lane l at 8k + 4h + l
  int carry[4];
  int pr, prm, ir, jr, is, is_old, level;
};

class Random {
 public:
  explicit Random(const RandomArg&);
  ~Random();
  Float Uniform();
  Float Gauss(Float);
  int Z(int);
  void ReadState(const char*);
  void WriteState(const char*);

 private:
  Random(const Random&);
  Random& operator=(const Random&);
  UniformDeviate* r;
  RandomType random_type;
};

// ------------------------------------------------------------------ lattice field
class Lattice {
 public:
  explicit Lattice(const LatticeArg&, Random*);
  ~Lattice();
  void Refresh();
  void Refresh(Float);                // then add a real shift to every site
  void Refresh(std::complex<Float>);  // then add a complex shift
  void Refresh(Float*);               // then add a site-dependent shift Float[2V]
  Float* Get();  // device pointer to Float[2V] (re, im)
  int FieldSize() { return field_size; }
  std::complex<Float> Mean();
  Float Variance();
  Float& operator[](std::size_t i) { return field[i]; }

 private:
  Lattice(const Lattice&);
  Lattice& operator=(const Lattice&);
  Random* rng_p;
  FieldType field_type;
  int field_size;
  std::vector<Float> field;
  Float* dev_field;
};

// ------------------------------------------------------------------ Fourier
class Fourier {
 public:
  explicit Fourier(int batch);
  ~Fourier();
  int BatchSize() { return batch_size; }
  void Forward();   // sign -1, k -> x
  void Backward();  // sign +1, x -> k
  fftComplex* Get(int n);

 private:
  Fourier(const Fourier&);
  Fourier& operator=(const Fourier&);
  int batch_size;
  fftComplex* b;
  nrc_fft* plan;
};

// ------------------------------------------------------------------ dispersion
class Dispersion {
 public:
  explicit Dispersion(DispersionType, Float mass, CutoffType);
  Float Get(int n1, int n2, int n3);
  Float Get(int i) { return dispersion[i]; }

 private:
  std::vector<Float> dispersion;
};

// ------------------------------------------------------------------ Hamiltonians
class Hamiltonian {
 public:
  Hamiltonian() {}
  virtual ~Hamiltonian() {}
  virtual void Evolve(Fourier*) = 0;
  virtual void Initialize() = 0;

 private:
  Hamiltonian(const Hamiltonian&);
  Hamiltonian& operator=(const Hamiltonian&);
};

class Kinetic : public Hamiltonian {
 public:
  explicit Kinetic(const KineticArg&, Float dt);
  ~Kinetic();
  void Evolve(Fourier*);
  void Initialize() {}
  const std::vector<Float>& Xi() const { return xi; }

 private:
  std::vector<Float> xi;
  Float* dev_xi;
};

class Interaction : public Hamiltonian {
 public:
  explicit Interaction(Lattice*, const InteractionArg&, Float dt);
  ~Interaction();
  void Evolve(Fourier*);
  void Initialize();
  const std::vector<Float>& FormFactor() const { return interaction; }

 private:
  Lattice* lattice;
  Fourier* field;
  std::vector<Float> interaction;
  Float* dev_interaction;
};

class Potential : public Hamiltonian {
 public:
  explicit Potential(const PotentialArg&, Float dt);
  ~Potential();
  void Evolve(Fourier*);
  void Initialize() {}

 private:
  std::vector<Float> potential;
  Float* dev_potential;
};

class Interactions : public Hamiltonian {
 public:
  explicit Interactions(std::vector<Hamiltonian*> list) : interactions(list) {}
  void Evolve(Fourier*) {}
  void Initialize();

 private:
  std::vector<Hamiltonian*> interactions;
};

// ------------------------------------------------------------------ Propagator
class Propagator {
 public:
  explicit Propagator(const PropagatorArg&);
  ~Propagator();
  void Run(std::vector<Hamiltonian*>);
  void Set(Float*, int);
  Float* Get(int);
  Fourier* Get() { return b; }
  void Normalize();
  void Normalize(Float);

 private:
  Propagator(const Propagator&);
  Propagator& operator=(const Propagator&);
  Fourier* b;
  int vol;
};

// ------------------------------------------------------------------ one-body sources / sinks
class OneBody {
 public:
  explicit OneBody(const OneBodyArg&);
  ~OneBody();
  void Set(std::string);         // lines "k1 k2 k3 re im" of a momentum-space source file
  void Set();                    // applies source_type (momentum sources: nothing to do)
  void Set(int, int, int);       // unit vector at momentum (k1, k2, k3)
  void Set(int, int, int, int);  // +k +- (-k) combination
  void Set(Float*);              // host vector Float[2V]
  void Set(const Vector3&);
  std::complex<Float> Project(Float*);  // <psi|phi>, phi a device pointer
  Float* Get();                         // device pointer

 private:
  OneBody(const OneBody&);
  OneBody& operator=(const OneBody&);
  void Upload();
  SourceType source_type;
  Float lambda1, lambda2, lambda3;
  std::vector<Float> psi;
  Float* dev_psi;
  double* dev_m;
  int vol;
};

// ------------------------------------------------------------------ special functions
// Dawson's integral (src/utils/special_functions.C:48-78): Rybicki's sampling-theorem
// formula, single-precision accurate (~2e-7), as the reference evaluates it.
Float DawsonF(Float x);
// Orthonormal Hermite polynomials h_n (int h_m h_n e^{-x^2} dx = delta_mn) by the three-term
// recurrence from h_0 = pi^(-1/4) (special_functions.C:7-34): h_n(x), and h[0..n-1](x).
Float HermiteH(int n, Float x);
void HermiteH(Float* h, int n, Float x);

// ------------------------------------------------------------------ pair sink
// The two-particle sink <pair| of SlaterDet / SlaterDet2 / SlaterDet3 (include/two_body.h:
// 11-36): the wave function w(p) and its pair matrix element sum_p w(p) G1(p) G2(-(p + bc)).
class TwoBody {
 public:
  explicit TwoBody(const TwoBodyArg&);
  ~TwoBody();
  void Deform(std::string);  // adds "k1 k2 k3 re" lines of a file to w
  std::complex<Float> Run(Float*, Float*);  // device pointers
  // phi[t + n s] = Run(props1[s], props2[t]) for s, t < n (device pointers)
  void Run(Float** props1, Float** props2, std::complex<Float>* phi, int n);

 private:
  TwoBody(const TwoBody&);
  TwoBody& operator=(const TwoBody&);
  void Upload();
  int vol;
  std::vector<double> wavefunc;
  std::vector<int> opposite_parity_index;
  double* dev_w;
  int* dev_opp;
};

// ------------------------------------------------------------------ Slater determinants
class SlaterDet {
 public:
  explicit SlaterDet(Float** props, OneBody** one_body, int n);
  explicit SlaterDet(Float** props1, Float** props2, TwoBody* two_body, int n);
  ~SlaterDet();
  std::complex<Float> Run(int n);
  std::complex<Float> SlaterMat(int s, int t) { return slater_mat[t + mat_size * s]; }
  void SetDetType(DetType t) { det_type = t; }

 private:
  SlaterDet(const SlaterDet&);
  SlaterDet& operator=(const SlaterDet&);
  std::vector<std::complex<Float> > slater_mat;
  int mat_size;
  DetType det_type;
};

// A pair-sink Slater matrix phi[t + N s] = <pair | props1_s props2_t> with the optional
// one-body matrices psi1[t + N s] = <one_body1_s | props1_t> and psi2 likewise, and the
// determinant of the leading n x n minor with row s0 (from psi2) or column t0 (from psi1)
// struck out and replaced (src/slater_det/slater_det2.C; slater_det3.C).  SlaterDet2 fills
// the matrices in its constructor, SlaterDet3 in Initialize(two_body).
class SlaterDet2 {
 public:
  explicit SlaterDet2(Float** props1, Float** props2, OneBody** one_body1, OneBody** one_body2,
                      TwoBody* two_body, int n);
  ~SlaterDet2();
  std::complex<Float> Run(int n, int s0, int t0);
  std::complex<Float> SlaterMat(int s, int t) { return phi[t + N * s]; }
  void SetDetType(DetType t) { det_type = t; }

 protected:
  SlaterDet2(int n);
  void Fill(Float** props1, Float** props2, OneBody** one_body1, OneBody** one_body2,
            TwoBody* two_body);
  int N;
  DetType det_type;
  std::vector<std::complex<Float> > phi, psi1, psi2;
  bool has_psi1, has_psi2;

 private:
  SlaterDet2(const SlaterDet2&);
  SlaterDet2& operator=(const SlaterDet2&);
};

class SlaterDet3 : public SlaterDet2 {
 public:
  explicit SlaterDet3(Float** props1, Float** props2, OneBody** one_body1, OneBody** one_body2,
                      int n);
  void Initialize(TwoBody* two_body);

 private:
  std::vector<Float*> props1, props2;
  std::vector<OneBody*> one_body1, one_body2;
};

// ------------------------------------------------------------------ LU (ludcmp.h)
void ludcmp(std::complex<Float>* a, int n, int* indx, Float* d);

// ------------------------------------------------------------------ momenta
class Momenta {
 public:
  explicit Momenta(const MomentaArg&);
  int MomentaCount() { return momenta_count; }
  std::vector<int> GetMomentum(int);
  int OppositeParityIndex(int n) { return parity_list[n]; }
  int ShellCount(int);
  int AccShellCount(int);
  int NumShells() { return (int)shell_count.size(); }

 private:
  std::vector<Float> omega;
  std::vector<std::vector<int> > momentum;
  std::vector<int> shell_count;
  std::vector<int> parity_list;
  int momenta_count;
};

#endif
