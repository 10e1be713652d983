// physics.cpp -- Random/Ranlxs, Lattice, Fourier, Dispersion, Kinetic, Interaction,
// Potential, Propagator, OneBody, SlaterDet, ludcmp and Momenta of the host class layer.
// Device work goes through the C ABI (nrcps.h); every status is checked with ERR.Check,
// which terminates like the reference's fatal errors.
//
// Reference behaviour: random.C:12-181, ranlxs.C:49-265, lattice.C:15-109, fourier.C:6-101
// (FFTW conventions) / fourier.cu:9-79, dispersion.C:11-93, hamiltonian.C:77-281,
// hamiltonian_evolve.cu:18-104, hamiltonian_initialize.cu:17-39, propagator.C:18-136,
// one_body.C:21-316, one_body_project.cu:17-34, slater_det.C:12-128, ludcmp.C:9-65,
// momenta.C:12-190.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <limits>

#include "../../include/nrcps/nrcps_host.h"

static nrc_context* ctx() { return nrc_host_context(); }

template <typename X>
static X* dev_alloc(size_t n, const char* fname) {
  void* p = nullptr;
  ERR.Check(fname, nrc_malloc(ctx(), &p, n * sizeof(X)));
  return (X*)p;
}

template <typename X>
static void upload(X* dst, const std::vector<X>& src, const char* fname) {
  ERR.Check(fname, nrc_memcpy_h2d(ctx(), dst, src.data(), src.size() * sizeof(X)));
}

// ------------------------------------------------------------------ Ranlxs
// Luscher's ranlxs: four interleaved 24-bit subtract-with-borrow lanes x_n = x_{n-10} -
// x_{n-24} - c, with pr pair-steps per refill of 96 outputs (pr = 109, 202, 397 for luxury
// 0, 1, 2).  State layout and seeding as the reference so that sequences and the 105-long
// state vector agree.
static const int RLX_BASE = 0x1000000, RLX_MASK = 0xffffff;

Ranlxs::Ranlxs(int lv) : pr(0), prm(0), ir(0), jr(7), is(95), is_old(0), level(lv) {
  if (lv < 0 || lv > 2) ERR.General("Ranlxs::Ranlxs(int)", "Bad luxury level (0, 1 or 2).");
  Init(1);
}

void Ranlxs::Init(long seed) {
  if (seed <= 0 || seed > 2147483647L)
    ERR.General("Ranlxs::Init(long)", "Bad choice of seed (should be between 1 and 2^31-1).");
  pr = level == 0 ? 109 : level == 1 ? 202 : 397;
  int bits[31];
  long s = seed;
  for (int k = 0; k < 31; ++k) {
    bits[k] = (int)(s % 2);
    s /= 2;
  }
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
      num[4 * k + lane] = (k % 4 == lane) ? RLX_MASK - v : v;
    }
  for (int l = 0; l < 4; ++l) carry[l] = 0;
  ir = 0;
  jr = 7;
  is = 95;
  is_old = 0;
  prm = pr % 12;
}

void Ranlxs::Update() {
  int i = ir, j = jr;
  for (int step = 0; step < pr; ++step) {
    int* xi = num + 8 * i;
    const int* xj = num + 8 * j;
    for (int l = 0; l < 4; ++l) {
      int d = xj[l] - xi[l] - carry[l];
      if (d < 0) {
        xi[4 + l] += 1;
        d += RLX_BASE;
      }
      xi[l] = d & RLX_MASK;
    }
    for (int l = 0; l < 4; ++l) {
      const int d = xj[4 + l] - xi[4 + l];
      carry[l] = d < 0;
      xi[4 + l] = (d + RLX_BASE) & RLX_MASK;
    }
    i = (i + 1) % 12;
    j = (j + 1) % 12;
  }
  ir = (ir + prm) % 12;
  jr = (jr + prm) % 12;
  is = 8 * ir;
  is_old = is;
}

Float Ranlxs::Run() {
  is = (is + 1) % 96;
  if (is == is_old) Update();
  return (Float)((float)std::ldexp(1.0, -24) * (float)num[is]);
}

void Ranlxs::GetState(long* st) {
  st[0] = 105;
  for (int k = 0; k < 96; ++k) st[k + 1] = num[k];
  for (int l = 0; l < 4; ++l) st[97 + l] = carry[l];
  st[101] = pr;
  st[102] = ir;
  st[103] = jr;
  st[104] = is;
}

void Ranlxs::SetState(long* st) {
  const char* fname = "Ranlxs::SetState(long*)";
  if (st[0] != 105) ERR.General(fname, "Unexpected input data.");
  for (int k = 0; k < 96; ++k)
    if (st[k + 1] < 0 || st[k + 1] > RLX_MASK) ERR.General(fname, "Unexpected input data.");
  for (int l = 0; l < 4; ++l)
    if (st[97 + l] != 0 && st[97 + l] != 1) ERR.General(fname, "Unexpected input data.");
  if ((st[101] != 109 && st[101] != 202 && st[101] != 397) || st[102] < 0 || st[102] > 11 ||
      st[103] != (st[102] + 7) % 12 || st[104] < 0 || st[104] > 95)
    ERR.General(fname, "Unexpected input data.");
  for (int k = 0; k < 96; ++k) num[k] = (int)st[k + 1];
  for (int l = 0; l < 4; ++l) carry[l] = (int)st[97 + l];
  pr = (int)st[101];
  ir = (int)st[102];
  jr = (int)st[103];
  is = (int)st[104];
  is_old = 8 * ir;
  prm = pr % 12;
  level = pr == 109 ? 0 : pr == 202 ? 1 : 2;
}

RandomType Ranlxs::TypeQ() {
  return level == 0 ? RANDOM_TYPE_RANLXS0 : level == 1 ? RANDOM_TYPE_RANLXS1 : RANDOM_TYPE_RANLXS2;
}

// ------------------------------------------------------------------ Random
Random::Random(const RandomArg& a) : r(nullptr), random_type(a.random_type) {
  const char* fname = "Random::Random(const RandomArg &)";
  switch (a.random_type) {
    case RANDOM_TYPE_RANLXS0: r = new Ranlxs(0); break;
    case RANDOM_TYPE_RANLXS1: r = new Ranlxs(1); break;
    case RANDOM_TYPE_RANLXS2: r = new Ranlxs(2); break;
    default: ERR.NotImplemented(fname, "random_type (only RANLXS0/1/2 are built).");
  }
  long seed = 1;
  switch (a.seed_type) {
    case SEED_TYPE_DEFAULT: seed = 1; break;
    case SEED_TYPE_INPUT:
      seed = a.seed;
      if (seed <= 0) ERR.General(fname, "Input seed must be positive.");
      break;
    case SEED_TYPE_TIME: seed = (long)std::time(nullptr); break;
    case SEED_TYPE_FILE: break;
  }
  if (a.seed_type == SEED_TYPE_FILE) {
    ReadState(a.file_stem.c_str());
  } else {
    seed += 491 * Comms::Rank();  // random.C:88
    r->Init(seed);
  }
}

Random::~Random() { delete r; }
Float Random::Uniform() { return r->Run(); }

Float Random::Gauss(Float sigma) {  // polar Box-Muller (random.C:99-111)
  Float x, y, r2;
  do {
    x = 2 * Uniform() - 1;
    y = 2 * Uniform() - 1;
    r2 = x * x + y * y;
  } while (r2 > 1.0 || r2 == 0);
  return sigma * y * std::sqrt(-2.0 * std::log(r2) / r2);
}

int Random::Z(int p) { return (int)std::floor(p * Uniform()); }

// <stem>.<rank>: RandomType (int) then long[StateSize] (random.C:134-181)
void Random::WriteState(const char* stem) {
  const char* fname = "Random::WriteState(const char*)";
  std::vector<long> st(r->StateSize());
  r->GetState(st.data());
  FILE* f = Fopen(IO_TYPE_ALL, stem, "wb");
  if (Fwrite(&random_type, sizeof(RandomType), 1, f) != 1)
    ERR.General(fname, "Cannot write seed type to file.");
  if ((int)Fwrite(st.data(), sizeof(long), st.size(), f) != (int)st.size())
    ERR.General(fname, "Cannot write state to file.");
  Fclose(f);
}

void Random::ReadState(const char* stem) {
  const char* fname = "Random::ReadState(const char*)";
  std::vector<long> st(r->StateSize());
  RandomType t;
  FILE* f = Fopen(IO_TYPE_ALL, stem, "rb");
  if (Fread(&t, sizeof(RandomType), 1, f) != 1) ERR.General(fname, "Cannot read seed type.");
  if (t != random_type) ERR.General(fname, "Random type of file does not match.");
  if ((int)Fread(st.data(), sizeof(long), st.size(), f) != (int)st.size())
    ERR.General(fname, "File may be corrupted.");
  Fclose(f);
  r->SetState(st.data());
}

// ------------------------------------------------------------------ Lattice
Lattice::Lattice(const LatticeArg& a, Random* rng)
    : rng_p(rng), field_type(a.field_type), field_size(2 * GJP.Vol()), field(field_size),
      dev_field(dev_alloc<Float>(field_size, "Lattice::Lattice")) {
  Refresh();
}

Lattice::~Lattice() { nrc_free(ctx(), dev_field); }

// lattice.C:60-109: host draws in site order, then one copy to the device
void Lattice::Refresh() {
  if (field_type == FIELD_TYPE_GAUSS) {
    for (int i = 0; i < field_size; i += 2) {
      field[i] = rng_p->Gauss(1.0);
      field[i + 1] = 0.0;
    }
  } else if (field_type == FIELD_TYPE_COMPLEXGAUSS) {
    for (int i = 0; i < field_size; i += 2) {
      field[i] = rng_p->Gauss(1.0);
      field[i + 1] = rng_p->Gauss(1.0);
    }
  } else if (field_type == FIELD_TYPE_ZTWO) {
    for (int i = 0; i < field_size; i += 2) {
      field[i] = 2.0 * rng_p->Z(2) - 1.0;
      field[i + 1] = 0.0;
    }
  } else {
    for (int i = 0; i < field_size; i += 2) {
      const int z = rng_p->Z(3);
      field[i] = z == 0 ? 1.0 : -0.5;
      field[i + 1] = z == 0 ? 0.0 : (z == 1 ? 0.8660254037844388 : -0.8660254037844388);
    }
  }
  upload(dev_field, field, "void Lattice::Refresh()");
}

// shifted fields for importance sampling (lattice.C:111-165; the reference's GPU build
// rejects these, here the shift is applied before the copy to the device)
void Lattice::Refresh(Float shift) { Refresh(std::complex<Float>(shift, 0.0)); }

void Lattice::Refresh(std::complex<Float> shift) {
  std::vector<Float> s((size_t)field_size);
  for (int i = 0; i < field_size; i += 2) {
    s[i] = shift.real();
    s[i + 1] = shift.imag();
  }
  Refresh(s.data());
}

void Lattice::Refresh(Float* shift) {
  Refresh();
  for (int i = 0; i < field_size; ++i) field[i] += shift[i];
  upload(dev_field, field, "void Lattice::Refresh(Float*)");
}

Float* Lattice::Get() { return dev_field; }

std::complex<Float> Lattice::Mean() {
  Float re = 0.0, im = 0.0;
  const int vol = field_size / 2;
  for (int i = 0; i < vol; ++i) {
    re += field[2 * i];
    im += field[2 * i + 1];
  }
  return std::complex<Float>(re / vol, im / vol);
}

Float Lattice::Variance() {
  const std::complex<Float> m = Mean();
  Float v = 0.0;
  for (int i = 0; i < field_size; ++i) v += field[i] * field[i];
  return v / (field_size / 2) - std::norm(m);
}

// ------------------------------------------------------------------ Fourier
Fourier::Fourier(int batch) : batch_size(batch), b(nullptr), plan(nullptr) {
  const char* fname = "Fourier::Fourier(int)";
  b = dev_alloc<fftComplex>((size_t)batch * GJP.Vol(), fname);
  ERR.Check(fname, nrc_fft_plan(ctx(), GJP.Xsites(), GJP.Ysites(), GJP.Zsites(), batch,
                                NRC_FLOAT_PREC, &plan));
}

Fourier::~Fourier() {
  nrc_fft_destroy(plan);
  nrc_free(ctx(), b);
}

void Fourier::Forward() { ERR.Check("void Fourier::Forward()", nrc_fft_exec(plan, b, -1)); }
void Fourier::Backward() { ERR.Check("void Fourier::Backward()", nrc_fft_exec(plan, b, +1)); }

fftComplex* Fourier::Get(int n) {
  const char* fname = "fftComplex* Fourier::Get(int)";
  if (n < 0) ERR.General(fname, "Argument must be non-negative.");
  if (n >= batch_size) ERR.General(fname, "Argument must be less than batch_size.");
  return b + (size_t)n * GJP.Vol();
}

// ------------------------------------------------------------------ Dispersion
Dispersion::Dispersion(DispersionType type, Float mass, CutoffType cutoff) : dispersion(GJP.Vol()) {
  Float bc[3];
  GJP.GetBcShift(bc);
  const int X = GJP.Xsites(), Y = GJP.Ysites(), Z = GJP.Zsites();
  const int bx = GJP.Xbc(), by = GJP.Ybc(), bz = GJP.Zbc();
  for (int i = 0; i < GJP.Vol(); ++i) {
    int x = i / (Y * Z), y = (i / Z) % Y, z = i % Z;
    if (type == DISPERSION_TYPE_STANDARD) {
      // 2 sum_j sin^2(p_j / 2) / mass
      const Float sx = std::sin((x + bc[0]) * PI / X);
      const Float sy = std::sin((y + bc[1]) * PI / Y);
      const Float sz = std::sin((z + bc[2]) * PI / Z);
      Float e = 2.0 * (sx * sx + sy * sy + sz * sz);
      e /= mass;
      dispersion[i] = e;
    } else {
      if (x >= X / 2) x = X - x - bx;
      if (y >= Y / 2) y = Y - y - by;
      if (z >= Z / 2) z = Z - z - bz;
      const Float sx = (x + bc[0]) / (Float)X, sy = (y + bc[1]) / (Float)Y, sz = (z + bc[2]) / (Float)Z;
      const Float psq = (sx * sx + sy * sy + sz * sz) * 4.0 * PISQ;
      if (cutoff == CUTOFF_TYPE_HARD && psq > GJP.CutoffSq())
        dispersion[i] = std::numeric_limits<Float>::infinity();
      else if (type == DISPERSION_TYPE_QUADRATIC)
        dispersion[i] = psq / (2.0 * mass);
      else
        dispersion[i] = std::exp(psq / (2.0 * mass)) - 1.0;
    }
  }
}

Float Dispersion::Get(int n1, int n2, int n3) {
  return dispersion[n3 + n2 * GJP.Zsites() + n1 * GJP.Ysites() * GJP.Zsites()];
}

// ------------------------------------------------------------------ Kinetic
Kinetic::Kinetic(const KineticArg& a, Float dt) : xi(GJP.Vol()) {
  Dispersion d(a.dispersion_type, a.mass / dt, CUTOFF_TYPE_HARD);
  for (int i = 0; i < GJP.Vol(); ++i) xi[i] = 1.0 + d.Get(i);
  dev_xi = dev_alloc<Float>(xi.size(), "Kinetic::Kinetic");
  upload(dev_xi, xi, "Kinetic::Kinetic");
}

Kinetic::~Kinetic() { nrc_free(ctx(), dev_xi); }

void Kinetic::Evolve(Fourier* b) {  // psi(k) /= xi(k) for every batch member
  ERR.Check("void Kinetic::Evolve(Fourier*)",
            nrc_scale_real(ctx(), b->Get(0), dev_xi, b->BatchSize(), GJP.Vol(), 1, NRC_FLOAT_PREC));
}

// ------------------------------------------------------------------ Interaction
Interaction::Interaction(Lattice* lat, const InteractionArg& a, Float dt)
    : lattice(lat), field(nullptr), interaction(GJP.Vol()) {
  const char* fname = "Interaction::Interaction(Lattice*, const InteractionArg &, Float)";
  const int vol = GJP.Vol();
  const Float mass = a.mass;
  Dispersion d(DISPERSION_TYPE_PERFECT, mass, CUTOFF_TYPE_NONE);
  const Float XI = std::exp(GJP.CutoffSq() / (2.0 * mass));
  for (int i = 0; i < vol; ++i) {
    Float xi = 1.0 + d.Get(i);
    if (xi > XI) xi = XI;
    Float psq = 0.0;
    if (a.interaction_type == INTERACTION_TYPE_ONEMINUSXIINVSQ) psq = 1.0 - 1.0 / (xi * xi);
    else if (a.interaction_type == INTERACTION_TYPE_XISQMINUSONE) psq = xi * xi - 1.0;
    else ERR.NotImplemented(fname, "Interaction type.");
    psq *= mass;
    Float v = 0.0;  // Taylor series in p^2
    for (int j = 0; j < a.num_couplings; ++j) v += a.couplings[j] * std::pow(psq, j);
    if (v < 0.0) ERR.General(fname, "Interaction is less the zero; cannot take square root.");
    v *= 2.0 * TWOPI / mass;
    v = std::sqrt(v);
    v /= vol;  // the FFTs are unnormalised
    v *= dt;
    interaction[i] = v;
  }
  dev_interaction = dev_alloc<Float>(vol, fname);
  upload(dev_interaction, interaction, fname);
  field = new Fourier(1);
}

Interaction::~Interaction() {
  delete field;
  nrc_free(ctx(), dev_interaction);
}

// field = (1 + F-[f F+[phi]]) / V (hamiltonian_initialize.cu:17-35)
void Interaction::Initialize() {
  const char* fname = "void Interaction::Initialize()";
  const int vol = GJP.Vol();
  ERR.Check(fname, nrc_memcpy_d2d(ctx(), field->Get(0), lattice->Get(), 2 * vol * sizeof(Float)));
  field->Backward();
  ERR.Check(fname, nrc_scale_real(ctx(), field->Get(0), dev_interaction, 1, vol, 0, NRC_FLOAT_PREC));
  field->Forward();
  ERR.Check(fname, nrc_plus_one_div_vol(ctx(), field->Get(0), vol, NRC_FLOAT_PREC));
}

void Interaction::Evolve(Fourier* b) {
  const char* fname = "void Interaction::Evolve(Fourier*)";
  if (GJP.APBCQuery()) ERR.NotImplemented(fname, "APBCs no allowed.");
  b->Forward();  // to position space
  ERR.Check(fname, nrc_mul_complex(ctx(), b->Get(0), field->Get(0), b->BatchSize(), GJP.Vol(),
                                   NRC_FLOAT_PREC));
  b->Backward();  // to momentum space
}

void Interactions::Initialize() {
  for (size_t n = 0; n < interactions.size(); ++n) interactions[n]->Initialize();
}

// ------------------------------------------------------------------ Potential
Potential::Potential(const PotentialArg& a, Float dt) : potential(GJP.Vol()) {
  const char* fname = "Potential::Potential(const PotentialArg &, Float)";
  if (GJP.APBCQuery()) ERR.NotImplemented(fname, "Must use PBCs when a potential is present.");
  const int X = GJP.Xsites(), Y = GJP.Ysites(), Z = GJP.Zsites(), vol = GJP.Vol();
  for (int i = 0; i < vol; ++i) {
    int z = i % Z, y = (i / Z) % Y, x = i / (Y * Z);
    if (x >= X / 2) x -= X;
    if (y >= Y / 2) y -= Y;
    if (z >= Z / 2) z -= Z;
    Float v = 0.0;
    if (a.potential_form == POTENTIAL_FORM_HARMONIC) {
      v = a.spring_constant1 * x * x;
      v += a.spring_constant2 * y * y;
      v += a.spring_constant3 * z * z;
      v /= 2.0;
    } else if (a.potential_form != POTENTIAL_FORM_NONE) {
      ERR.NotImplemented(fname, "Potential form not supported.");
    }
    Float p = a.potential_type == POTENTIAL_TYPE_LIN ? 1.0 - dt * v : std::exp(-dt * v);
    p /= vol;
    potential[i] = p;
  }
  dev_potential = dev_alloc<Float>(vol, fname);
  upload(dev_potential, potential, fname);
}

Potential::~Potential() { nrc_free(ctx(), dev_potential); }

void Potential::Evolve(Fourier* b) {
  const char* fname = "void Potential::Evolve(Fourier*)";
  b->Forward();
  ERR.Check(fname, nrc_scale_real(ctx(), b->Get(0), dev_potential, b->BatchSize(), GJP.Vol(), 0,
                                  NRC_FLOAT_PREC));
  b->Backward();
}

// ------------------------------------------------------------------ Propagator
Propagator::Propagator(const PropagatorArg& a) : b(new Fourier(a.n_fermions)), vol(GJP.Vol()) {}
Propagator::~Propagator() { delete b; }

void Propagator::Run(std::vector<Hamiltonian*> hs) {
  for (size_t i = 0; i < hs.size(); ++i) hs[i]->Evolve(b);
}

void Propagator::Set(Float* psi, int n) {
  ERR.Check("void Propagator::Set(Float*, int)",
            nrc_memcpy_d2d(ctx(), Get(n), psi, 2 * (size_t)vol * sizeof(Float)));
}

Float* Propagator::Get(int n) { return (Float*)b->Get(n); }

// propagator.C:78-110: every orbital to unit norm (host-only in the reference; here on the
// device: norms with FP64 sums, then one division per orbital)
void Propagator::Normalize() {
  const char* fname = "void Propagator::Normalize()";
  const int n = b->BatchSize();
  std::vector<double> norms(n);
  ERR.Check(fname, nrc_norms(ctx(), b->Get(0), n, vol, NRC_FLOAT_PREC, norms.data()));
  for (int k = 0; k < n; ++k)
    ERR.Check(fname, nrc_scale_const(ctx(), b->Get(k), vol, norms[k], 1, NRC_FLOAT_PREC));
}

// propagator.C:112-133: every orbital divided by `norm`
void Propagator::Normalize(Float norm) {
  ERR.Check("void Propagator::Normalize(Float)",
            nrc_scale_const(ctx(), b->Get(0), (long)b->BatchSize() * vol, norm, 1,
                            NRC_FLOAT_PREC));
}

// ------------------------------------------------------------------ OneBody
OneBody::OneBody(const OneBodyArg& a)
    : source_type(a.source_type), lambda1(a.lambda1), lambda2(a.lambda2), lambda3(a.lambda3),
      psi(2 * GJP.Vol(), 0.0), vol(GJP.Vol()) {
  dev_psi = dev_alloc<Float>(psi.size(), "OneBody::OneBody");
  dev_m = dev_alloc<double>(2, "OneBody::OneBody");
}

OneBody::~OneBody() {
  nrc_free(ctx(), dev_psi);
  nrc_free(ctx(), dev_m);
}

void OneBody::Upload() { upload(dev_psi, psi, "OneBody::Set"); }

// SOURCE_TYPE_SHO (one_body.C:53-171): psi holds the coefficients c(n1, n2, n3) at momentum
// index (n1, n2, n3), n < min(10, L); the source becomes, at each lattice momentum p_i =
// 2 pi k_i / L (k centred on 0),
//   sum_n c(n) h_n1(d1) h_n2(d2) h_n3(d3) exp(-(d1^2 + d2^2 + d3^2) / 2)
//       sqrt(lambda1 lambda2 lambda3) sqrt((2 pi)^3 / V),
// d1 = p1 lambda1, d2 = p2 lambda2, d3 = p3 lambda2 (the reference scales the third axis by
// lambda2, kept so that sources agree with its files).
void OneBody::Set() {
  const char* fname = "void OneBody::Set()";
  if (source_type == SOURCE_TYPE_SHO) {
    if (GJP.APBCQuery()) ERR.NotImplemented(fname, "Must use PBCs when a potential is present.");
    const int X = GJP.Xsites(), Y = GJP.Ysites(), Z = GJP.Zsites();
    const int n1m = X < 10 ? X : 10, n2m = Y < 10 ? Y : 10, n3m = Z < 10 ? Z : 10;
    VRB.Warn(fname, "Maximum SHO levels (nx,ny,nz) are (%d,%d,%d)", n1m, n2m, n3m);
    std::vector<double> c(2 * (size_t)n1m * n2m * n3m);
    for (int a = 0; a < n1m; ++a)
      for (int b = 0; b < n2m; ++b)
        for (int d = 0; d < n3m; ++d) {
          const int j = 2 * Vector3ToIndex(Vector3(a, b, d));
          c[2 * ((a * n2m + b) * n3m + d)] = psi[j];
          c[2 * ((a * n2m + b) * n3m + d) + 1] = psi[j + 1];
        }
    std::vector<Float> h1(n1m), h2(n2m), h3(n3m);
    const double norm = std::sqrt(lambda1 * lambda2 * lambda3) *
                        std::sqrt(TWOPI * TWOPI * TWOPI / vol);
    for (int i = 0; i < vol; ++i) {
      const Vector3 v = IndexToVector3(i);
      int x = v.GetX(), y = v.GetY(), z = v.GetZ();
      if (x >= X / 2) x -= X;
      if (y >= Y / 2) y -= Y;
      if (z >= Z / 2) z -= Z;
      const double dx = TWOPI * x * lambda1 / X, dy = TWOPI * y * lambda2 / Y,
                   dz = TWOPI * z * lambda2 / Z;
      HermiteH(h1.data(), n1m, dx);
      HermiteH(h2.data(), n2m, dy);
      HermiteH(h3.data(), n3m, dz);
      double re = 0.0, im = 0.0;
      for (int a = 0; a < n1m; ++a)
        for (int b = 0; b < n2m; ++b)
          for (int d = 0; d < n3m; ++d) {
            const double t = h1[a] * h2[b] * h3[d];
            re += c[2 * ((a * n2m + b) * n3m + d)] * t;
            im += c[2 * ((a * n2m + b) * n3m + d) + 1] * t;
          }
      const double g = std::exp(-(dx * dx + dy * dy + dz * dz) / 2.0) * norm;
      psi[2 * i] = re * g;
      psi[2 * i + 1] = im * g;
    }
  }
  Upload();
}

// one_body.C:172-212: coefficients of momentum basis vectors, one "k1 k2 k3 re im" per line
void OneBody::Set(std::string name) {
  const char* fname = "void OneBody::Set(string)";
  std::fill(psi.begin(), psi.end(), 0.0);
  std::ifstream file(name.c_str());
  if (!file) ERR.FileR(fname, name.c_str());
  int k1, k2, k3;
  Float re, im;
  while (file >> k1 >> k2 >> k3 >> re >> im) {
    const int j = 2 * Vector3ToIndex(Vector3(k1, k2, k3));
    psi[j] = re;
    psi[j + 1] = im;
  }
  Set();
}

// one_body.C:215-252: the unit vector (and its parity partner), then Set() as the reference
// does, so that an SHO-typed source set by momentum is the SHO transform of those coefficients
void OneBody::Set(int k1, int k2, int k3) {
  std::fill(psi.begin(), psi.end(), 0.0);
  psi[2 * Vector3ToIndex(Vector3(k1, k2, k3))] = 1.0;
  Set();
}

void OneBody::Set(int k1, int k2, int k3, int parity) {
  std::fill(psi.begin(), psi.end(), 0.0);
  psi[2 * Vector3ToIndex(Vector3(k1, k2, k3))] = 1.0;
  psi[2 * Vector3ToIndex(-1 * Vector3(k1, k2, k3))] = parity ? 1.0 : -1.0;
  Set();
}

void OneBody::Set(const Vector3& v) { Set(v.GetX(), v.GetY(), v.GetZ()); }

void OneBody::Set(Float* src) {
  for (int i = 0; i < 2 * vol; ++i) psi[i] = src[i];
  Upload();
}

Float* OneBody::Get() { return dev_psi; }

std::complex<Float> OneBody::Project(Float* phi) {
  const char* fname = "complex<Float> OneBody::Project(Float*)";
  const int mode = NRC_FLOAT_PREC == NRC_PREC_F64 ? NRC_OVERLAP_F64 : NRC_OVERLAP_F32;
  ERR.Check(fname, nrc_overlap(ctx(), dev_psi, phi, dev_m, 1, 1, vol, 1, mode));
  double m[2];
  ERR.Check(fname, nrc_memcpy_d2h(ctx(), m, dev_m, sizeof(m)));
  return std::complex<Float>((Float)m[0], (Float)m[1]);
}

// ------------------------------------------------------------------ SlaterDet
// M[t + n s] = <sink_s | orbital_t> (slater_det.C:25-39).  When the orbitals are one
// propagator's contiguous batch, each sink row is one batched overlap launch and the matrix
// crosses to the host once; LU and determinants run on the device (nrc_lu_det) on
// contiguous storage.
SlaterDet::SlaterDet(Float** props, OneBody** one_body, int n)
    : slater_mat((size_t)n * n), mat_size(n), det_type(DET_TYPE_STANDARD) {
  const char* fname = "SlaterDet::SlaterDet(Float**, OneBody**, int)";
  const long vol = GJP.Vol();
  bool contiguous = true;
  for (int t = 1; t < n; ++t) contiguous = contiguous && props[t] == props[0] + 2 * vol * t;
  const int mode = NRC_FLOAT_PREC == NRC_PREC_F64 ? NRC_OVERLAP_F64 : NRC_OVERLAP_F32;
  double* dev_M = dev_alloc<double>(2 * (size_t)n * n, fname);
  for (int s = 0; s < n; ++s) {
    if (contiguous) {
      ERR.Check(fname, nrc_overlap(ctx(), one_body[s]->Get(), props[0], dev_M + 2 * (size_t)s * n,
                                   1, n, vol, 1, mode));
    } else {
      for (int t = 0; t < n; ++t)
        ERR.Check(fname, nrc_overlap(ctx(), one_body[s]->Get(), props[t],
                                     dev_M + 2 * ((size_t)s * n + t), 1, 1, vol, 1, mode));
    }
  }
  std::vector<double> h(2 * (size_t)n * n);
  ERR.Check(fname, nrc_memcpy_d2h(ctx(), h.data(), dev_M, h.size() * sizeof(double)));
  nrc_free(ctx(), dev_M);
  for (size_t i = 0; i < (size_t)n * n; ++i)
    slater_mat[i] = std::complex<Float>((Float)h[2 * i], (Float)h[2 * i + 1]);
}

SlaterDet::~SlaterDet() {}

// det (STANDARD) or log det (LOG) of the n x n matrix h[s][t] (interleaved re, im) by the
// device LU; a singular matrix ends the program as ludcmp does (ludcmp.C:23)
static std::complex<Float> det_of(const std::vector<double>& h, int n, DetType det_type,
                                  const char* fname) {
  double* dev = dev_alloc<double>(h.size() + 4 + 1, fname);
  int* dev_st = dev_alloc<int>(1, fname);
  ERR.Check(fname, nrc_memcpy_h2d(ctx(), dev, h.data(), h.size() * sizeof(double)));
  double* out = dev + h.size();  // logabs, phase (2), argsum
  ERR.Check(fname, nrc_lu_det(ctx(), dev, n, n, 1, 0, out, out + 1, out + 3, dev_st));
  double r[4];
  int st = 0;
  ERR.Check(fname, nrc_memcpy_d2h(ctx(), r, out, sizeof(r)));
  ERR.Check(fname, nrc_memcpy_d2h(ctx(), &st, dev_st, sizeof(int)));
  nrc_free(ctx(), dev);
  nrc_free(ctx(), dev_st);
  if (st) {
    std::printf("ludcmp(complex<Float>*, int, int*, Float*): singular matrix\n");
    std::exit(EXIT_FAILURE);
  }
  if (det_type == DET_TYPE_LOG) return std::complex<Float>((Float)r[0], (Float)r[3]);
  const double m = std::exp(r[0]);
  return std::complex<Float>((Float)(m * r[1]), (Float)(m * r[2]));
}

// det (STANDARD) or log det (LOG) of the leading n x n minor (slater_det.C:59-115)
std::complex<Float> SlaterDet::Run(int n) {
  const char* fname = "complex<Float> SlaterDet::Run(int)";
  if (n > mat_size) ERR.General(fname, "Submatrix exceeds size of matrix.");
  if (n < 1) {
    VRB.Warn(fname, "n < 1, returning unity.");
    return 1.0;
  }
  std::vector<double> h(2 * (size_t)n * n);
  for (int s = 0; s < n; ++s)
    for (int t = 0; t < n; ++t) {
      h[2 * ((size_t)s * n + t)] = slater_mat[t + mat_size * s].real();
      h[2 * ((size_t)s * n + t) + 1] = slater_mat[t + mat_size * s].imag();
    }
  return det_of(h, n, det_type, fname);
}

// M(s, t) = <pair| props1_s props2_t> (slater_det.C:41-57)
SlaterDet::SlaterDet(Float** props1, Float** props2, TwoBody* two_body, int n)
    : slater_mat((size_t)n * n), mat_size(n), det_type(DET_TYPE_STANDARD) {
  two_body->Run(props1, props2, slater_mat.data(), n);
}

// ------------------------------------------------------------------ special functions
// Rybicki (1989): F(x) = lim_{h->0} pi^-1/2 sum over odd n of exp(-(x - n h)^2) / n, summed
// around the even n0 nearest x/h with h = 0.4 and six odd terms each side; below 0.2 the
// Taylor series to x^7.  Step, term count and evaluation order are special_functions.C's
// (48-78), so the wave functions built on it agree with the reference's to rounding.
Float DawsonF(Float x) {
  const double h = 0.4;
  if (std::fabs(x) < 0.2) {
    const double x2 = x * x;
    return x * (1.0 - (2.0 / 3.0) * x2 * (1.0 - 0.4 * x2 * (1.0 - (2.0 / 7.0) * x2)));
  }
  double c[7];
  for (int i = 1; i <= 6; ++i) c[i] = std::exp(-((2.0 * i - 1.0) * h) * ((2.0 * i - 1.0) * h));
  const double ax = std::fabs(x);
  const int n0 = 2 * (int)(0.5 * ax / h + 0.5);
  const double xp = ax - n0 * h;
  double e1 = std::exp(2.0 * xp * h);
  const double e2 = e1 * e1;
  double d1 = n0 + 1, d2 = d1 - 2.0, sum = 0.0;
  for (int i = 1; i <= 6; ++i, d1 += 2.0, d2 -= 2.0, e1 *= e2)
    sum += c[i] * (e1 / d1 + 1.0 / (d2 * e1));
  const double m = std::exp(-xp * xp);
  return (Float)((x >= 0.0 ? m : -m) * sum / 1.7724538509055160273);
}

// orthonormal Hermite polynomials: h_0 = pi^(-1/4), h_j = x sqrt(2/j) h_{j-1} -
// sqrt((j-1)/j) h_{j-2}
Float HermiteH(int n, Float x) {
  Float p1 = PIM4, p2 = 0.0;
  for (int j = 1; j <= n; ++j) {
    const Float p3 = p2;
    p2 = p1;
    p1 = x * std::sqrt(2.0 / j) * p2 - std::sqrt((Float)(j - 1) / j) * p3;
  }
  return p1;
}

void HermiteH(Float* h, int n, Float x) {
  Float a = 0.0;
  h[0] = PIM4;
  for (int j = 1; j < n; ++j) {
    const Float b = a;
    a = h[j - 1];
    h[j] = x * std::sqrt(2.0 / j) * a - std::sqrt((Float)(j - 1) / j) * b;
  }
}

// ------------------------------------------------------------------ TwoBody
// w(p) of the two-body wave function types (two_body.C:21-128) and the site of -(p + bc)
// (two_body.C:31-33); the pair matrix elements run on the device (nrc_pair_overlap).
TwoBody::TwoBody(const TwoBodyArg& a)
    : vol(GJP.Vol()), wavefunc(GJP.Vol(), 0.0), opposite_parity_index(GJP.Vol()) {
  const char* fname = "TwoBody::TwoBody(const TwoBodyArg &)";
  const int bx = GJP.Xbc(), by = GJP.Ybc(), bz = GJP.Zbc();
  for (int i = 0; i < vol; ++i) {
    const Vector3 v = IndexToVector3(i);
    opposite_parity_index[i] =
        Vector3ToIndex(Vector3(-(v.GetX() + bx), -(v.GetY() + by), -(v.GetZ() + bz)));
  }
  if (a.wavefunc_type == WAVEFUNC_TYPE_UNIFORM) {
    for (int i = 0; i < vol; ++i) wavefunc[i] = 1.0 / vol;
  } else if (a.wavefunc_type == WAVEFUNC_TYPE_GND) {
    Dispersion d1(a.dispersion_type1, a.mass1, CUTOFF_TYPE_HARD);
    Dispersion d2(a.dispersion_type2, a.mass2, CUTOFF_TYPE_HARD);
    for (int i = 0; i < vol; ++i) {
      const double xi1 = 1.0 + d1.Get(i), xi2 = 1.0 + d2.Get(i);
      double w = 1.0;
      w /= a.lambda - 1.0 / (xi1 * xi2);  // finite where xi is infinite (HARD cutoff)
      wavefunc[i] = w;
    }
  } else if (a.wavefunc_type != WAVEFUNC_TYPE_NONE) {
    Dispersion d(DISPERSION_TYPE_QUADRATIC, 1.0, CUTOFF_TYPE_HARD);
    for (int i = 0; i < vol; ++i) {
      const double psq = 2.0 * d.Get(i);
      double w = 0.0;
      if (a.wavefunc_type == WAVEFUNC_TYPE_TGND) {
        if (psq <= GJP.CutoffSq()) {
          const double xi = std::sqrt(psq) / (2.0 * a.lambda);
          w = xi < 1e-15 ? 1.0 : DawsonF(xi) / xi;
        }
      } else if (a.wavefunc_type == WAVEFUNC_TYPE_PAIR1) {
        w = 1.0 / (psq + a.lambda);
      } else {
        w = psq < 1e-15 ? 0.0 : std::exp(-a.lambda * std::sqrt(psq)) / psq;
      }
      wavefunc[i] = w;
    }
  }
  dev_w = dev_alloc<double>(vol, fname);
  dev_opp = dev_alloc<int>(vol, fname);
  ERR.Check(fname, nrc_memcpy_h2d(ctx(), dev_opp, opposite_parity_index.data(),
                                  vol * sizeof(int)));
  Upload();
}

TwoBody::~TwoBody() {
  nrc_free(ctx(), dev_w);
  nrc_free(ctx(), dev_opp);
}

void TwoBody::Upload() {
  ERR.Check("TwoBody::Upload", nrc_memcpy_h2d(ctx(), dev_w, wavefunc.data(), vol * sizeof(double)));
}

// two_body.C:145-196: each "k1 k2 k3 re" line adds re to w at that momentum
void TwoBody::Deform(std::string name) {
  const char* fname = "void TwoBody::Deform(string)";
  std::ifstream file(name.c_str());
  if (!file) ERR.FileR(fname, name.c_str());
  int k1, k2, k3;
  Float re;
  while (file >> k1 >> k2 >> k3 >> re) {
    VRB.Flow(fname, "Appending %f to %d component of wave function.", (double)re,
             Vector3ToIndex(Vector3(k1, k2, k3)));
    wavefunc[Vector3ToIndex(Vector3(k1, k2, k3))] += re;
  }
  Upload();
}

static int pair_mode() {
  return NRC_FLOAT_PREC == NRC_PREC_F64 ? NRC_OVERLAP_F64 : NRC_OVERLAP_F32;
}

std::complex<Float> TwoBody::Run(Float* p1, Float* p2) {
  const char* fname = "complex<Float> TwoBody::Run(Float*, Float*)";
  double* dev_m = dev_alloc<double>(2, fname);
  ERR.Check(fname, nrc_pair_overlap(ctx(), p1, p2, dev_w, dev_opp, dev_m, 1, 1, vol, 1,
                                    pair_mode()));
  double m[2];
  ERR.Check(fname, nrc_memcpy_d2h(ctx(), m, dev_m, sizeof(m)));
  nrc_free(ctx(), dev_m);
  return std::complex<Float>((Float)m[0], (Float)m[1]);
}

void TwoBody::Run(Float** props1, Float** props2, std::complex<Float>* phi, int n) {
  const char* fname = "void TwoBody::Run(Float**, Float**, complex<Float>*, int)";
  bool contiguous = true;
  for (int t = 1; t < n; ++t)
    contiguous = contiguous && props1[t] == props1[0] + 2 * (long)vol * t &&
                 props2[t] == props2[0] + 2 * (long)vol * t;
  double* dev_M = dev_alloc<double>(2 * (size_t)n * n, fname);
  if (contiguous) {
    ERR.Check(fname, nrc_pair_overlap(ctx(), props1[0], props2[0], dev_w, dev_opp, dev_M, n, n,
                                      vol, 1, pair_mode()));
  } else {
    for (int s = 0; s < n; ++s)
      for (int t = 0; t < n; ++t)
        ERR.Check(fname, nrc_pair_overlap(ctx(), props1[s], props2[t], dev_w, dev_opp,
                                          dev_M + 2 * ((size_t)s * n + t), 1, 1, vol, 1,
                                          pair_mode()));
  }
  std::vector<double> h(2 * (size_t)n * n);
  ERR.Check(fname, nrc_memcpy_d2h(ctx(), h.data(), dev_M, h.size() * sizeof(double)));
  nrc_free(ctx(), dev_M);
  for (size_t i = 0; i < (size_t)n * n; ++i)
    phi[i] = std::complex<Float>((Float)h[2 * i], (Float)h[2 * i + 1]);
}

// ------------------------------------------------------------------ SlaterDet2 / SlaterDet3
// one-body matrix out[t + n s] = <one_body_s | props_t>
static void one_body_matrix(Float** props, OneBody** one_body, int n,
                            std::vector<std::complex<Float> >& out) {
  SlaterDet m(props, one_body, n);
  for (int s = 0; s < n; ++s)
    for (int t = 0; t < n; ++t) out[t + n * s] = m.SlaterMat(s, t);
}

SlaterDet2::SlaterDet2(int n)
    : N(n), det_type(DET_TYPE_STANDARD), phi((size_t)n * n), psi1((size_t)n * n),
      psi2((size_t)n * n), has_psi1(false), has_psi2(false) {}

// slater_det2.C:25-60
SlaterDet2::SlaterDet2(Float** props1, Float** props2, OneBody** one_body1,
                       OneBody** one_body2, TwoBody* two_body, int n)
    : SlaterDet2(n) {
  Fill(props1, props2, one_body1, one_body2, two_body);
}

SlaterDet2::~SlaterDet2() {}

void SlaterDet2::Fill(Float** props1, Float** props2, OneBody** one_body1, OneBody** one_body2,
                      TwoBody* two_body) {
  has_psi1 = one_body1 != nullptr;
  has_psi2 = one_body2 != nullptr;
  if (has_psi1) one_body_matrix(props1, one_body1, N, psi1);
  if (has_psi2) one_body_matrix(props2, one_body2, N, psi2);
  two_body->Run(props1, props2, phi.data(), N);
}

// slater_det2.C:78-150: the leading n x n minor of phi with column t0 replaced by
// psi1[s + t0 N] or row s0 by psi2[t + s0 N] (at most one of them non-negative)
std::complex<Float> SlaterDet2::Run(int n, int s0, int t0) {
  const char* fname = "complex<Float> SlaterDet2::Run(int, int, int)";
  if (n > N) ERR.General(fname, "Sub-matrix exceeds size of matrix.");
  if (n < 1) ERR.General(fname, "Specified size of sub-matrix is not positive.");
  if (s0 >= 0 && t0 >= 0) ERR.General(fname, "s0 and t0 cannot both be non-negative.");
  if (t0 >= 0 && t0 < n && !has_psi1)
    ERR.General(fname, "t0 cannot be non-negative; psi1==NULL.");
  if (s0 >= 0 && s0 < n && !has_psi2)
    ERR.General(fname, "s0 cannot be non-negative; psi2==NULL.");
  std::vector<double> h(2 * (size_t)n * n);
  for (int s = 0; s < n; ++s)
    for (int t = 0; t < n; ++t) {
      std::complex<Float> v = phi[t + N * s];
      if (t == t0) v = psi1[s + t0 * N];
      if (s == s0) v = psi2[t + s0 * N];
      h[2 * ((size_t)s * n + t)] = v.real();
      h[2 * ((size_t)s * n + t) + 1] = v.imag();
    }
  return det_of(h, n, det_type, fname);
}

// slater_det3.C:18-75: the propagators and sinks are kept; Initialize(two_body) recomputes
// the matrices (slater_det3.C:95-125), so one object serves every time slice
SlaterDet3::SlaterDet3(Float** p1, Float** p2, OneBody** ob1, OneBody** ob2, int n)
    : SlaterDet2(n), props1(p1, p1 + n), props2(p2, p2 + n) {
  if (ob1) one_body1.assign(ob1, ob1 + n);
  if (ob2) one_body2.assign(ob2, ob2 + n);
}

void SlaterDet3::Initialize(TwoBody* two_body) {
  Fill(props1.data(), props2.data(), one_body1.empty() ? nullptr : one_body1.data(),
       one_body2.empty() ? nullptr : one_body2.data(), two_body);
}

// ------------------------------------------------------------------ ludcmp
// The reference's host routine, run by the device kernel that reproduces it bit for bit.
void ludcmp(std::complex<Float>* a, int n, int* indx, Float* d) {
  const char* fname = "ludcmp(complex<Float>*, int, int*, Float*)";
  std::vector<double> h(2 * (size_t)n * n);
  for (size_t i = 0; i < (size_t)n * n; ++i) {
    h[2 * i] = a[i].real();
    h[2 * i + 1] = a[i].imag();
  }
  double* dev = dev_alloc<double>(h.size() + 1, fname);
  int* di = dev_alloc<int>((size_t)n + 1, fname);
  ERR.Check(fname, nrc_memcpy_h2d(ctx(), dev, h.data(), h.size() * sizeof(double)));
  ERR.Check(fname, nrc_ludcmp(ctx(), dev, n, 1, di, dev + h.size(), di + n));
  std::vector<int> hi((size_t)n + 1);
  double dd = 1.0;
  ERR.Check(fname, nrc_memcpy_d2h(ctx(), h.data(), dev, h.size() * sizeof(double)));
  ERR.Check(fname, nrc_memcpy_d2h(ctx(), hi.data(), di, hi.size() * sizeof(int)));
  ERR.Check(fname, nrc_memcpy_d2h(ctx(), &dd, dev + h.size(), sizeof(double)));
  nrc_free(ctx(), dev);
  nrc_free(ctx(), di);
  if (hi[n]) {
    std::printf("ludcmp(complex<Float>*, int, int*, Float*): singular matrix\n");
    std::exit(EXIT_FAILURE);
  }
  for (size_t i = 0; i < (size_t)n * n; ++i)
    a[i] = std::complex<Float>((Float)h[2 * i], (Float)h[2 * i + 1]);
  for (int i = 0; i < n; ++i) indx[i] = hi[i];
  *d = (Float)dd;
}

// ------------------------------------------------------------------ Momenta
Momenta::Momenta(const MomentaArg& a) : momenta_count(0) {
  const char* fname = "Momenta::Momenta(const MomentaArg &)";
  const int X = GJP.Xsites(), Y = GJP.Ysites(), Z = GJP.Zsites();
  Dispersion d(a.dispersion_type, a.mass, CUTOFF_TYPE_HARD);
  for (int n1 = 0; n1 < X; ++n1)
    for (int n2 = 0; n2 < Y; ++n2)
      for (int n3 = 0; n3 < Z; ++n3) {
        const Float e = std::log(1 + d.Get(n1, n2, n3));
        if (e <= a.fermi_energy) {
          omega.push_back(e);
          momentum.push_back(std::vector<int>{n1, n2, n3});
        }
      }
  momenta_count = (int)omega.size();
  if (!momenta_count) ERR.General(fname, "No energy levels exceed the specified fermi level.");
  // insertion sort by energy, with the reference's lower bound j > 0 (momenta.C:62-73)
  for (int i = 1; i < momenta_count; ++i) {
    const Float e = omega[i];
    const std::vector<int> tmp = momentum[i];
    int j = i - 1;
    while (j > 0 && omega[j] > e) {
      omega[j + 1] = omega[j];
      momentum[j + 1] = momentum[j];
      --j;
    }
    omega[j + 1] = e;
    momentum[j + 1] = tmp;
  }
  shell_count.push_back(1);
  for (int i = 1; i < momenta_count; ++i) {
    if (omega[i] == omega[i - 1]) shell_count.back()++;
    else shell_count.push_back(1);
  }
  parity_list.assign(momenta_count, 0);
  const int dims[3] = {X, Y, Z};
  const int bcs[3] = {GJP.Xbc(), GJP.Ybc(), GJP.Zbc()};
  for (int i = 0; i < momenta_count; ++i) {
    std::vector<int> p(3);
    for (int ax = 0; ax < 3; ++ax) p[ax] = (dims[ax] - momentum[i][ax] - bcs[ax]) % dims[ax];
    bool found = false;
    for (int j = 0; j < momenta_count; ++j)
      if (p == momentum[j]) {
        parity_list[i] = j;
        found = true;
      }
    if (!found) ERR.General(fname, "Opposite parity vector is missing; try adjusting the fermi energy.");
  }
}

std::vector<int> Momenta::GetMomentum(int n) {
  if (n >= momenta_count || n < 0)
    ERR.General("vector<int> Momenta::GetMomentum(int)", "Specified momentum exceeds momentum count.");
  return momentum[n];
}

int Momenta::ShellCount(int n) {
  if (n < 0 || n >= NumShells()) ERR.General("Momenta::ShellCount(int)", "Invalid argument.");
  return shell_count[n];
}

int Momenta::AccShellCount(int n) {
  if (n < 0) return 0;
  if (n >= NumShells()) ERR.General("Momenta::AccShellCount(int)", "Invalid argument.");
  int acc = 0;
  for (int s = 0; s <= n; ++s) acc += shell_count[s];
  return acc;
}
