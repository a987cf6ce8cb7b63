// Host side of the source-level drop-in (include/classicaldft/classicaldft.h).
// Every evaluation and solver step goes through the C ABI of <cdft.h>.
#include "classicaldft/classicaldft.h"

#include <algorithm>
#include <cstring>
#include <cstdio>
#include <random>
#include <fstream>
#include <limits>
#include <sstream>
#include <functional>

namespace {

void check(int rc) {
  if (rc == CDFT_OK) return;
  if (rc == CDFT_ETA_TOO_LARGE) throw Eta_Too_Large_Exception();
  throw std::runtime_error(std::string("cdft: ") + cdft_last_error());
}

// Neumaier accumulator (include/Summation.h:15-56)
struct NSum {
  double s = 0, c = 0;
  void add(double v) {
    const double t = s + v;
    if (std::fabs(s) >= std::fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
  }
  double sum() const { return s + c; }
};

}  // namespace

// ------------------------------------------------------------------ DFT_Vec
void DFT_Vec::IncrementBy(const DFT_Vec& o) {
  for (size_t i = 0; i < d_.size(); ++i) d_[i] += o.d_[i];
}
void DFT_Vec::IncrementBy_Scaled_Vector(const DFT_Vec& o, double s) {
  for (size_t i = 0; i < d_.size(); ++i) d_[i] += o.d_[i] * s;
}
void DFT_Vec::MultBy(double s) {
  for (double& v : d_) v *= s;
}
void DFT_Vec::add(double s) {
  for (double& v : d_) v += s;
}
double DFT_Vec::inf_norm() const {
  double m = 0;
  for (double v : d_) m = std::max(m, std::fabs(v));
  return m;
}
double DFT_Vec::euclidean_norm() const {
  double s = 0;
  for (double v : d_) s += v * v;
  return std::sqrt(s);
}
double DFT_Vec::dotWith(const DFT_Vec& o) const {
  double s = 0;
  for (size_t i = 0; i < d_.size(); ++i) s += d_[i] * o.d_[i];
  return s;
}
double DFT_Vec::accu() const {
  double s = 0;
  for (double v : d_) s += v;
  return s;
}
double DFT_Vec::min() const { return *std::min_element(d_.begin(), d_.end()); }
double DFT_Vec::max() const { return *std::max_element(d_.begin(), d_.end()); }

// ------------------------------------------------------------------ Lattice
// Lattice::init, include/Lattice.h:57-85
Lattice::Lattice(double dx, double dy, double dz, double L[]) : dx_(dx), dy_(dy), dz_(dz) {
  Nx_ = long((L[0] + 1e-8 * dx_) / dx_);
  Ny_ = long((L[1] + 1e-8 * dy_) / dy_);
  Nz_ = long((L[2] + 1e-8 * dz_) / dz_);
  L_[0] = Nx_ * dx_;
  L_[1] = Ny_ * dy_;
  L_[2] = Nz_ * dz_;
  if (std::fabs(L[0] - Nx_ * dx_) > 1e-5 * dx_) throw std::runtime_error("Box length incommensurate with lattice in x direction");
  if (std::fabs(L[1] - Ny_ * dy_) > 1e-5 * dy_) throw std::runtime_error("Box length incommensurate with lattice in y direction");
  if (std::fabs(L[2] - Nz_ * dz_) > 1e-5 * dz_) throw std::runtime_error("Box length incommensurate with lattice in z direction");
}

long Lattice::pos(int i, int j, int k) const {
  if (i < 0 || i >= Nx_) throw std::runtime_error("ix out of bounds");
  if (j < 0 || j >= Ny_) throw std::runtime_error("iy out of bounds");
  if (k < 0 || k >= Nz_) throw std::runtime_error("iz out of bounds");
  return k + Nz_ * (j + Ny_ * i);
}

long Lattice::get_PBC_Pos(int ix, int iy, int iz) const {
  auto w = [](long v, long n) { v %= n; return v < 0 ? v + n : v; };
  return w(iz, Nz_) + Nz_ * (w(iy, Ny_) + Ny_ * w(ix, Nx_));
}

// ------------------------------------------------------------------ Density
Density::Density(double dx, double L[]) : Lattice(dx, L), rho_(Ntot()), vWall_(Ntot()) {}
Density::Density(double dx, double dy, double dz, double L[]) : Lattice(dx, dy, dz, L), rho_(Ntot()), vWall_(Ntot()) {}
void Density::set(double val) {
  for (long i = 0; i < rho_.size(); ++i) rho_.set(i, val);
}
double Density::getNumberAtoms() const {
  NSum s;
  for (long i = 0; i < rho_.size(); ++i) s.add(rho_.get(i));
  return s.sum() * dV();
}
bool Density::has_field_() const { return vWall_.inf_norm() > 0.0; }

// ------------------------------------------------------------------ Species
int Species::SequenceNumber_ = 0;
Species::Species(Density& density, double mu, int seq) : density_(density), dF_(density.Ntot()), mu_(mu) {
  if (seq >= 0) {
    seq_num_ = seq;
    SequenceNumber_ = seq + 1;
  } else {
    seq_num_ = SequenceNumber_++;
  }
}

void Species::beginForceCalculation() {
  if (fixedMass_ > 0.0) {
    density_.scale_to(fixedMass_ / density_.getNumberAtoms());
    mu_ = 0.0;
  }
}

double Species::endForceCalculation(double multiplier) {
  if (fixedBackground_ && fixedMass_ > 0.0)
    throw std::runtime_error("Cannot have both fixed background and fixed mass .... aborting");
  if (fixedMass_ > 0.0) mu_ = multiplier;
  return 0;
}

FMT_Species::FMT_Species(Density& density, double hsd, double mu, int seq) : Species(density, mu, seq), hsd_(hsd) {
  // src/FMT_Species.cpp:321-322
  if (hsd < std::min(std::min(density.getDX(), density.getDY()), density.getDZ()))
    throw std::runtime_error("hsd is less than the lattice spacing ... aborting");
}

namespace {
int ao_force_of(const Species* s) {
  auto* fs = dynamic_cast<const FMT_Species*>(s);
  return fs ? fs->ao_force() : CDFT_AO_FORCE_EXACT;
}
}  // namespace

// ---------------------------------------------------------------------- FMT
FMT::~FMT() {
  if (sys_) cdft_system_destroy(sys_);
  if (mix_) cdft_mixture_destroy(mix_);
}

std::string esFMT::Name() const {
  return "esFMT with A = " + std::to_string(A_) + " and B = " + std::to_string(B_);
}

cdft_system FMT::system_for(const Species& s) const {
  const Lattice& L = s.getLattice();
  if (sys_ && sys_lattice_ == &L && sys_hsd_ == s.getHSD() && sys_species_ == &s && sys_aof_ == ao_force_of(&s))
    return sys_;
  if (sys_) cdft_system_destroy(sys_);
  cdft_system_desc d{};
  d.Nx = (int)L.Nx();
  d.Ny = (int)L.Ny();
  d.Nz = (int)L.Nz();
  d.dx = L.getDX();
  d.dy = L.getDY();
  d.dz = L.getDZ();
  d.model = model_;
  d.A = A_;
  d.B = B_;
  d.hsd = s.getHSD();
  d.precision = precision_;
  if (auto* fs = dynamic_cast<const FMT_Species*>(&s)) {
    d.ao_Rp = fs->ao_Rp();
    d.ao_reservoir = fs->ao_reservoir();
    d.ao_force = fs->ao_force();
  }
  check(cdft_system_create(&d, &sys_));
  sys_lattice_ = &L;
  sys_hsd_ = s.getHSD();
  sys_aof_ = ao_force_of(&s);
  sys_species_ = &s;
  return sys_;
}

cdft_mixture FMT::mixture_for(const std::vector<FMT_Species*>& species) const {
  std::vector<const FMT_Species*> key(species.begin(), species.end());
  std::vector<double> hsd;
  for (const FMT_Species* fs : species) {
    if (fs->ao_Rp() > 0) throw std::runtime_error("an AO polymer is implemented for a single FMT species");
    const Lattice &L = fs->getLattice(), &L0 = species.front()->getLattice();
    if (L.Nx() != L0.Nx() || L.Ny() != L0.Ny() || L.Nz() != L0.Nz() || L.getDX() != L0.getDX() ||
        L.getDY() != L0.getDY() || L.getDZ() != L0.getDZ())
      throw std::runtime_error("the species of a mixture must share one lattice");
    hsd.push_back(fs->getHSD());
  }
  if (mix_ && key == mix_species_ && hsd == mix_hsd_) return mix_;
  if (mix_) cdft_mixture_destroy(mix_);
  mix_ = nullptr;
  const Lattice& L = species.front()->getLattice();
  cdft_mixture_desc d{};
  d.Nx = (int)L.Nx();
  d.Ny = (int)L.Ny();
  d.Nz = (int)L.Nz();
  d.dx = L.getDX();
  d.dy = L.getDY();
  d.dz = L.getDZ();
  d.model = model_;
  d.A = A_;
  d.B = B_;
  d.nspecies = (int)hsd.size();
  d.hsd = hsd.data();
  check(cdft_mixture_create(&d, &mix_));
  mix_species_ = key;
  mix_hsd_ = hsd;
  return mix_;
}

// src/FMT.cpp:363-427: one FMT species runs the single-species pipeline, several the mixture's (the measures sum
// over species, FMT::getWeightedDensities src/FMT.cpp:78-111).  Both add dF_hs to each species' force.
double FMT::calculateFreeEnergyAndDerivatives(std::vector<Species*>& allSpecies) {
  std::vector<FMT_Species*> fmt;
  for (Species* s : allSpecies)
    if (auto* fs = dynamic_cast<FMT_Species*>(s)) fmt.push_back(fs);
  if (fmt.empty()) return 0.0;
  double F = 0.0;
  if (fmt.size() == 1) {
    cdft_system sys = system_for(*fmt[0]);
    check(cdft_fmt_eval(sys, fmt[0]->getDensity().get_density_pointer(), fmt[0]->getDF().memptr(), &F,
                        CDFT_ACCUMULATE));
    return F;
  }
  cdft_mixture mix = mixture_for(fmt);
  std::vector<const double*> rho;
  std::vector<double*> dF;
  for (FMT_Species* fs : fmt) {
    rho.push_back(fs->getDensity().get_density_pointer());
    dF.push_back(fs->getDF().memptr());
  }
  check(cdft_mixture_fmt_eval(mix, rho.data(), dF.data(), &F, CDFT_ACCUMULATE));
  return F;
}

void FMT::add_second_derivative(const std::vector<DFT_Vec>& v, std::vector<DFT_Vec>& d2F,
                                const std::vector<Species*>& allSpecies) {
  if (allSpecies.empty()) throw std::runtime_error("No species for FMT::add_second_derivative to work with");
  for (size_t i = 0; i < allSpecies.size(); ++i) {
    auto* fs = dynamic_cast<FMT_Species*>(allSpecies[i]);
    if (!fs) continue;
    if (i > 0) throw std::runtime_error("FMT on the GPU is implemented for a single species");
    check(cdft_fmt_hessian_dot(system_for(*fs), fs->getDensity().get_density_pointer(), v[i].memptr(),
                               d2F[i].memptr(), CDFT_ACCUMULATE));
  }
}

double FMT::BulkFex(const std::vector<double>& x, const std::vector<Species*>& allSpecies) const {
  for (size_t i = 0; i < allSpecies.size(); ++i)
    if (auto* fs = dynamic_cast<FMT_Species*>(allSpecies[i])) {
      double f, mu;
      check(cdft_system_bulk(system_for(*fs), x[i], &f, &mu));
      return f - (x[i] * std::log(x[i]) - x[i]);
    }
  return 0.0;
}

double FMT::BulkMuex(const std::vector<double>& x, const std::vector<Species*>& allSpecies, int species) const {
  auto* fs = dynamic_cast<FMT_Species*>(allSpecies[species]);
  if (!fs) return 0.0;
  double f, mu;
  check(cdft_system_bulk(system_for(*fs), x[species], &f, &mu));
  return mu - std::log(x[species]);
}

// --------------------------------------------------------------- Potential
cdft_potential_desc Potential1::desc() const {
  cdft_potential_desc p{};
  p.kind = kind();
  p.sigma = sigma_;
  p.eps = eps_;
  p.rcut = rcut_;
  p.alpha = alpha();
  p.bh = bh_ ? 1 : 0;
  p.r_att_min = r_att_min_;
  return p;
}

double Potential1::Watt(double r) const {
  const cdft_potential_desc p = desc();
  double w = 0.0;
  check(cdft_potential_watt(&p, 1, &r, &w));
  return w;
}

double Potential1::V0(double r) const {  // include/Potential1.h:77-81
  if (bh_) return r < r0_ ? V(r) : 0.0;
  return r < rmin_ ? V(r) - Vmin_ : 0.0;
}

double Potential1::getHSD(double kT) const {
  const cdft_potential_desc p = desc();
  double d = 0.0;
  check(cdft_potential_hsd(&p, kT, &d));
  return d;
}

double Potential1::getVDW_Parameter(double kT) const {
  const cdft_potential_desc p = desc();
  double a = 0.0;
  check(cdft_potential_vdw(&p, kT, &a));
  return a;
}

LJ::LJ(double sigma, double eps, double rcut) : Potential1(sigma, eps, rcut) {  // include/Potential1.h:140-145
  shift_ = rcut_ < 0 ? 0.0 : vr(rcut_);
  rmin_ = getRmin();
  Vmin_ = V(rmin_);
  r0_ = std::pow(0.5 * std::sqrt(1 + shift_) + 0.5, -1.0 / 6.0);
}

tWF::tWF(double sigma, double eps, double rcut, double alpha) : Potential1(sigma, eps, rcut), alpha_(alpha) {
  shift_ = rcut_ <= 0.0 ? 0.0 : vr(rcut_);  // include/Potential1.h:199-206
  rmin_ = getRmin();
  Vmin_ = V(rmin_);
  r0_ = std::sqrt(1 + std::pow(25 * std::sqrt(1 + shift_) + 25, -1.0 / 3.0));
}

double tWF::vr(double r) const {  // include/Potential1.h:234-243
  if (r < sigma_) return 1e50;
  const double s = r / sigma_, y = 1.0 / (s * s - 1), y3 = y * y * y;
  return (4 * eps_ / (alpha_ * alpha_)) * (y3 * y3 - alpha_ * y3);
}

WHDF::WHDF(double sigma, double eps, double rcut) : Potential1(sigma, eps, rcut) {  // include/Potential1.h:264-271
  eps_eff_ = eps * 2 * std::pow(rcut / sigma, 2) * std::pow(2 * ((rcut / sigma) * (rcut / sigma) - 1) / 3.0, -3.0);
  shift_ = 0.0;
  rmin_ = getRmin();
  Vmin_ = V(rmin_);
  r0_ = 1.0;
}

double WHDF::vr(double r) const {  // include/Potential1.h:298-303
  const double y = sigma_ / r, z = rcut_ / r;
  return eps_eff_ * (y * y - 1) * (z * z - 1) * (z * z - 1);
}

// ------------------------------------------------------------- Interaction
Interaction_Base::Interaction_Base(Species* s1, Species* s2, Potential1* v, double kT, int scheme, int ngauss)
    : s1_(s1), s2_(s2), v_(v), kT_(kT), scheme_(scheme), ngauss_(ngauss) {
  if (s1 != s2) throw std::runtime_error("only the single-species (s1 == s2) mean field is implemented");
}

void Interaction_Base::describe(cdft_system_desc& d) const {
  ensure_kernel();
  d.mf_kernel = kernel_.empty() ? nullptr : kernel_.data();
  d.mean_field = 1;
  d.lj_sigma = v_->getSigma();
  d.lj_eps = v_->getEps();
  d.lj_rcut = v_->getRcut();
  d.kT = kT_;
  d.mf_potential = v_->kind();
  d.mf_alpha = v_->alpha();
  d.mf_bh = v_->isBH() ? 1 : 0;
  d.mf_r_att_min = v_->isBH() ? 0.0 : v_->getWCALimit();
  d.mf_scheme = scheme_;
  d.mf_ngauss = ngauss_;
  d.mf_scale = scale_;
}

double Interaction_Base::getW(long s) const {
  if (w_.empty()) {
    const Lattice& L = s1_->getLattice();
    cdft_system_desc d{};
    d.Nx = (int)L.Nx();
    d.Ny = (int)L.Ny();
    d.Nz = (int)L.Nz();
    d.dx = L.getDX();
    d.dy = L.getDY();
    d.dz = L.getDZ();
    d.model = CDFT_MODEL_NONE;
    describe(d);
    if (d.mf_kernel) {
      w_.assign(d.mf_kernel, d.mf_kernel + L.Ntot());
      for (double& x : w_) x *= scale_;
    } else {
      cdft_system sys = nullptr;
      check(cdft_system_create(&d, &sys));
      w_.resize(L.Ntot());
      const int rc = cdft_system_get_mf_kernel(sys, w_.data());
      cdft_system_destroy(sys);
      check(rc);
    }
  }
  return w_.at(s);
}

double Interaction_Base::getVDWParameter() const {
  if (std::isnan(a_vdw_)) {
    const Lattice& L = s1_->getLattice();
    cdft_system_desc d{};
    d.Nx = (int)L.Nx();
    d.Ny = (int)L.Ny();
    d.Nz = (int)L.Nz();
    d.dx = L.getDX();
    d.dy = L.getDY();
    d.dz = L.getDZ();
    d.model = CDFT_MODEL_NONE;
    describe(d);
    cdft_system s = nullptr;
    check(cdft_system_create(&d, &s));
    check(cdft_system_a_vdw(s, &a_vdw_));
    cdft_system_destroy(s);
  }
  return a_vdw_;
}

std::string Potential1::getIdentifier() const {
  std::stringstream ss;
  const char* name = kind() == CDFT_POT_TWF ? "TWF_" : (kind() == CDFT_POT_WHDF ? "WHDF_" : "LJ_");
  ss << name << sigma_ << "_" << eps_ << "_";
  if (kind() == CDFT_POT_TWF) ss << alpha() << "_";
  ss << rcut_ << r_att_min_ << "_" << bh_;
  return ss.str();
}

Interaction::Interaction(Species* s1, Species* s2, Potential1* v, double kT, std::string pointsFile)
    : Interaction_Base(s1, s2, v, kT), pointsFile_(std::move(pointsFile)) {
  const Lattice& L = s1->getLattice();
  std::stringstream ss;
  ss << "weights_" << s1->getSequenceNumber() << "_" << s2->getSequenceNumber() << "_" << L.Nx() << "_" << L.Ny()
     << "_" << L.Nz() << "_" << v->getIdentifier() << "_" << ".dat";
  weightsFile_ = ss.str();
}

bool Interaction::readWeights() {
  const Lattice& L = s1_->getLattice();
  std::ifstream in(weightsFile_.c_str(), std::ios::binary);
  if (!in.good()) return false;
  std::string buf;
  std::getline(in, buf);
  std::stringstream s1(buf);
  long nx = 0, ny = 0, nz = 0;
  double dx = 0;
  s1 >> nx >> ny >> nz >> dx;
  bool ok = nx == L.Nx() && ny == L.Ny() && nz == L.Nz() && std::fabs(L.getDX() - dx) <= 1e-8 * (L.getDX() + dx);
  std::getline(in, buf);
  std::stringstream s2(buf);
  std::string identifier;
  s2 >> identifier;
  ok = ok && identifier == v_->getIdentifier();
  std::getline(in, buf);
  ok = ok && buf == pointsFile_;
  if (!ok) return false;
  long n = 0;
  in >> n;
  if (n != L.Ntot()) return false;
  kernel_.assign(n, 0.0);
  for (long i = 0; i < n; ++i) in >> kernel_[i];
  if (!in) return false;
  for (double& w : kernel_) w /= kT_;  // Interaction_Base::initialize: the temperature
  return true;
}

void Interaction::generateWeights() {
  const Lattice& L = s1_->getLattice();
  std::ifstream in(pointsFile_.c_str());
  if (!in.good()) throw std::runtime_error("File " + pointsFile_ + " could not be opened");
  std::vector<double> pts;
  for (std::string buf; std::getline(in, buf);) {
    std::stringstream st(buf);
    std::vector<double> line;
    for (double d; st >> d;) line.push_back(d);
    if (line.empty()) continue;
    if (line.size() < 3) throw std::runtime_error("a point of " + pointsFile_ + " needs three coordinates");
    line.push_back(4 * M_PI);
    const double r = std::sqrt(line[0] * line[0] + line[1] * line[1] + line[2] * line[2]);
    pts.insert(pts.end(), {line[0] / r, line[1] / r, line[2] / r, line[3]});
  }
  const long np = (long)pts.size() / 4;
  for (long i = 0; i < np; ++i) pts[4 * i + 3] /= np;
  std::vector<double> w(L.Ntot());
  const cdft_potential_desc pd = v_->desc();
  check(cdft_mf_kernel_points((int)L.Nx(), (int)L.Ny(), (int)L.Nz(), L.getDX(), L.getDY(), L.getDZ(), &pd, pts.data(),
                              (int)np, 0, w.data()));
  {  // the cache: header, identifier, points file, then the kernel (DFT_Vec::save, default stream precision)
    std::ofstream of(weightsFile_.c_str(), std::ios::binary);
    of.flags(std::ios::scientific);
    of.precision(std::numeric_limits<double>::digits10 + 1);
    of << L.Nx() << " " << L.Ny() << " " << L.Nz() << " " << L.getDX() << std::endl;
    of << v_->getIdentifier() << std::endl;
  }
  std::ofstream of(weightsFile_.c_str(), std::ios::binary | std::ios::app);
  of << pointsFile_ << std::endl;
  of << w.size() << std::endl;
  for (double x : w) of << x << " ";
  of << std::endl;
  kernel_.resize(w.size());
  for (size_t i = 0; i < w.size(); ++i) kernel_[i] = w[i] / kT_;
}

void Interaction::ensure_kernel() const {
  if (!kernel_.empty()) return;
  auto* self = const_cast<Interaction*>(this);
  self->from_cache_ = self->readWeights();
  if (!self->from_cache_) self->generateWeights();
}

// src/Interaction.cpp:61-72
double Interaction_Base::Mu(const std::vector<double>& x, int) const { return getVDWParameter() * x[0]; }
double Interaction_Base::Fhelmholtz(const std::vector<double>& x) const { return 0.5 * getVDWParameter() * x[0] * x[0]; }

// ---------------------------------------------------------------------- DFT
DFT::DFT(Species* s) {
  if (s) addSpecies(s);
}
DFT::~DFT() { invalidate(); }
void DFT::invalidate() const {
  if (sys_) cdft_system_destroy(sys_);
  sys_ = nullptr;
  if (mix_) cdft_mixture_destroy(mix_);
  mix_ = nullptr;
  mix_key_.clear();
}
void DFT::addSpecies(Species* s) {
  species_.push_back(s);
  invalidate();
}
void DFT::addInteraction(Interaction_Base* i) {
  if (!interactions_.empty()) throw std::runtime_error("one mean-field interaction is supported");
  interactions_.push_back(i);
  invalidate();
}
void DFT::addHardCoreContribution(FMT* fmt) {
  fmt_ = fmt;
  invalidate();
}
void DFT::set_precision(int p) {
  precision_ = p;
  invalidate();
}

cdft_mixture DFT::mixture() const {
  const Lattice& L0 = species_.front()->getLattice();
  std::vector<double> key, hsd, mu, mass;
  std::vector<int> bnd;
  std::vector<const double*> vext;
  for (const Species* s : species_) {
    const auto* fs = dynamic_cast<const FMT_Species*>(s);
    if (!fs || fs->ao_Rp() > 0)
      throw std::runtime_error("a DFT of several species takes hard-sphere FMT species (no AO polymer)");
    const Lattice& L = s->getLattice();
    if (L.Nx() != L0.Nx() || L.Ny() != L0.Ny() || L.Nz() != L0.Nz() || L.getDX() != L0.getDX() ||
        L.getDY() != L0.getDY() || L.getDZ() != L0.getDZ())
      throw std::runtime_error("the species of a mixture must share one lattice");
    const Density& D = s->getDensity();
    hsd.push_back(fs->getHSD());
    // with a fixed mass mu_ holds the last multiplier and the evaluation runs at mu = 0
    mu.push_back(s->is_mass_fixed() ? 0.0 : s->getChemPotential());
    mass.push_back(s->is_mass_fixed() ? s->getFixedMass() : 0.0);
    bnd.push_back((s->is_background_fixed() ? CDFT_FIXED_BACKGROUND : 0) |
                  (s->is_boundary_homogeneous() ? CDFT_HOMOGENEOUS_BOUNDARY : 0));
    vext.push_back(D.has_external_field() ? D.get_external_field().memptr() : nullptr);
    key.insert(key.end(), {hsd.back(), mu.back(), mass.back(), (double)bnd.back(), (double)D.field_version(),
                           (double)(size_t)s});
  }
  if (!fmt_) throw std::runtime_error("a DFT of several species needs a hard-core (FMT) contribution");
  if (!interactions_.empty())
    throw std::runtime_error("mean-field interactions are implemented for a single species");
  key.push_back(fmt_->model_id());
  if (mix_ && key == mix_key_) return mix_;
  if (mix_) cdft_mixture_destroy(mix_);
  mix_ = nullptr;
  cdft_mixture_desc d{};
  d.Nx = (int)L0.Nx();
  d.Ny = (int)L0.Ny();
  d.Nz = (int)L0.Nz();
  d.dx = L0.getDX();
  d.dy = L0.getDY();
  d.dz = L0.getDZ();
  d.model = fmt_->model_id();
  d.A = fmt_->A();
  d.B = fmt_->B();
  d.nspecies = (int)species_.size();
  d.hsd = hsd.data();
  d.ideal = 1;
  d.mu = mu.data();
  d.vext = vext.data();
  d.fixed_mass = mass.data();
  d.boundary = bnd.data();
  check(cdft_mixture_create(&d, &mix_));
  mix_key_ = key;
  return mix_;
}

cdft_system DFT::system() const {
  if (species_.size() != 1)
    throw std::runtime_error("this operation is implemented for a single species (several: DFT::mixture)");
  Species* s = species_.front();
  const Density& D = s->getDensity();
  // with a fixed mass mu_ holds the last multiplier and the evaluation runs at mu = 0
  const double mu = s->is_mass_fixed() ? 0.0 : s->getChemPotential();
  const double mass = s->is_mass_fixed() ? s->getFixedMass() : 0.0;
  const int bnd = (s->is_background_fixed() ? CDFT_FIXED_BACKGROUND : 0) |
                  (s->is_boundary_homogeneous() ? CDFT_HOMOGENEOUS_BOUNDARY : 0);
  const double mfs = interactions_.empty() ? 1.0 : interactions_.front()->scale();
  if (sys_ && sys_mu_ == mu && sys_mass_ == mass && sys_bnd_ == bnd && sys_mfs_ == mfs &&
      sys_field_ == D.field_version() && sys_aof_ == ao_force_of(s))
    return sys_;
  invalidate();
  cdft_system_desc d{};
  d.Nx = (int)D.Nx();
  d.Ny = (int)D.Ny();
  d.Nz = (int)D.Nz();
  d.dx = D.getDX();
  d.dy = D.getDY();
  d.dz = D.getDZ();
  d.model = fmt_ ? fmt_->model_id() : CDFT_MODEL_NONE;
  d.A = fmt_ ? fmt_->A() : 1.0;
  d.B = fmt_ ? fmt_->B() : 0.0;
  d.hsd = s->getHSD();
  d.precision = precision_;
  if (auto* fs = dynamic_cast<const FMT_Species*>(s)) {
    d.ao_Rp = fs->ao_Rp();
    d.ao_reservoir = fs->ao_reservoir();
    d.ao_force = fs->ao_force();
  }
  d.ideal = 1;
  d.mu = mu;
  d.fixed_mass = mass;
  d.boundary = bnd;
  d.vext = D.has_external_field() ? D.get_external_field().memptr() : nullptr;
  if (!interactions_.empty()) interactions_.front()->describe(d);
  check(cdft_system_create(&d, &sys_));
  sys_mu_ = mu;
  sys_mass_ = mass;
  sys_bnd_ = bnd;
  sys_mfs_ = mfs;
  sys_aof_ = ao_force_of(s);
  sys_field_ = D.field_version();
  return sys_;
}

// src/DFT.cpp:360-425
double DFT::calculateFreeEnergyAndDerivatives(bool onlyFex) {
  double parts[5];
  if (species_.size() > 1) {  // a hard-sphere mixture: every species' force written by the device
    cdft_mixture m = mixture();
    std::vector<const double*> r;
    std::vector<double*> f;
    for (Species* s : species_) {
      s->zeroForce();
      s->beginForceCalculation();
      r.push_back(s->getDensity().get_density_pointer());
      f.push_back(s->getDF().memptr());
    }
    check(cdft_mixture_eval(m, r.data(), f.data(), parts, onlyFex ? CDFT_NO_IDEAL : 0));
    std::vector<double> mu(species_.size());
    check(cdft_mixture_chem_potential(m, mu.data()));
    for (size_t i = 0; i < species_.size(); ++i) species_[i]->endForceCalculation(mu[i]);
    F_id_ = parts[1];
    F_hs_ = parts[2];
    F_mf_ = 0.0;
    F_ext_ = parts[4];
    return parts[0];
  }
  Species* s = species_.front();
  s->zeroForce();
  s->beginForceCalculation();
  cdft_system sys = system();
  // onlyFex skips the ideal gas only; the external field and mu stay (src/DFT.cpp:381-420)
  check(cdft_system_eval(sys, s->getDensity().get_density_pointer(), s->getDF().memptr(), parts,
                         onlyFex ? CDFT_NO_IDEAL : 0));
  F_id_ = parts[1];
  F_hs_ = parts[2];
  F_mf_ = parts[3];
  F_ext_ = parts[4];
  double mu = 0.0;
  check(cdft_system_chem_potential(sys, &mu));
  s->endForceCalculation(mu);
  return parts[0];
}

void Density::writeDensity(std::string filename) const {
  std::ofstream of(filename.c_str(), std::ios::binary);
  if (!of.good()) throw std::runtime_error("Could not open output file " + filename);
  of << rho_.size() << std::endl;
  for (long i = 0; i < rho_.size(); ++i) of << rho_.get(i) << " ";
  of << std::endl;
}

void Density::readDensity(const char* filename) {
  std::ifstream in(filename, std::ios::binary);
  if (!in.good()) {
    std::stringstream s;
    s << "Could not open input file " << filename << " ... aborting ... " << std::endl;
    throw std::runtime_error(s.str());
  }
  long n = -1;
  in >> n;
  if (n != rho_.size())
    throw std::runtime_error(std::string("Density file ") + filename + " holds " + std::to_string(n) +
                             " values, the lattice has " + std::to_string(rho_.size()));
  for (long i = 0; i < n; ++i) {
    double v;
    in >> v;
    rho_.set(i, v);
  }
  if (in.fail()) throw std::runtime_error(std::string("Could not read the values of ") + filename);
}

void DFT::matrix_dot_v_intern(const std::vector<DFT_Vec>& v, std::vector<DFT_Vec>& result, void*) const {
  Species* s = species_.front();
  const long N = s->getDensity().Ntot();
  if (v.size() < 1 || v[0].size() != N) throw std::runtime_error("DFT::matrix_dot_v: one direction of Ntot values");
  if (result.size() < 1) result.resize(1);
  if (result[0].size() != N) result[0].resize(N);
  check(cdft_system_hessian_dot(system(), s->getDensity().get_density_pointer(), v[0].memptr(), result[0].memptr(),
                                full_hessian_ ? 0 : CDFT_NO_IDEAL));
}

void DFT::matrix_dot_v(const std::vector<DFT_Vec>& v, std::vector<DFT_Vec>& result, void* param) const {
  matrix_dot_v_intern(v, result, param);
  if (use_squared_matrix_) {  // src/Dynamical_Matrix.cpp:38-49 (a DFT is not dynamic: no sign change)
    std::vector<DFT_Vec> v1(1);
    v1[0].set(result[0]);
    matrix_dot_v_intern(v1, result, param);
  }
}

// Zero the boundary points ix, iy or iz = 0 (Dynamical_Matrix::set_boundary_points_to_zero).
static void zero_boundary(const Lattice& L, DFT_Vec& v) {
  const long Nx = L.Nx(), Ny = L.Ny(), Nz = L.Nz();
  for (long ix = 0; ix < Nx; ++ix)
    for (long iy = 0; iy < Ny; ++iy)
      for (long iz = 0; iz < Nz; ++iz)
        if (ix == 0 || iy == 0 || iz == 0) v.set(iz + Nz * (iy + Ny * ix), 0.0);
}

// Dynamical_Matrix::matrix_dot_v1 (src/Dynamical_Matrix.cpp:52-70): with a fixed background the boundary entries
// of the input are zeroed before the product and those of the result after it.
void DFT::matrix_dot_v1(const DFT_Vec& v, DFT_Vec& result, void* param) const {
  const Species* s = species_.front();
  std::vector<DFT_Vec> vw(1), rw(1);
  vw[0].set(v);
  if (s->is_fixed_boundary()) zero_boundary(s->getLattice(), vw[0]);
  matrix_dot_v(vw, rw, param);
  result.set(rw[0]);
  if (s->is_fixed_boundary()) zero_boundary(s->getLattice(), result);
}

void DDFT::matrix_dot_v1(const DFT_Vec& v, DFT_Vec& result, void*) const {
  const Density& d = dft_->getDensity(0);
  if (v.size() != d.Ntot()) throw std::runtime_error("DDFT::matrix_dot_v: one direction of Ntot values");
  if (result.size() != d.Ntot()) result.zeros(d.Ntot());
  check(cdft_system_dynamic_dot(dft_->system(), d.get_density_pointer(), v.memptr(), result.memptr(),
                                (use_squared_matrix_ ? CDFT_SQUARED_MATRIX : 0) |
                                    (dft_->full_hessian_ ? 0 : CDFT_NO_IDEAL) |
                                    (central_ ? CDFT_CENTRAL_DIFFERENCES : 0)));
}

Log_Det::Log_Det(const DDFT& matrix, int order, double lam_max, double lam_min, bool verbose)
    : matrix_(matrix.getDFT()), order_(order), lam_max_(lam_max), lam_min_(lam_min), verbose_(verbose),
      dynamic_(true), ddft_(&matrix) {}

Eigenvalues::Eigenvalues(const DDFT& matrix, bool verbose) : matrix_(matrix.getDFT()), verbose_(verbose),
                                                              dynamic_(true), ddft_(&matrix) {}

// Flags of the matrix a Log_Det / Eigenvalues acts on: the DFT's Hessian with its own full_hessian_ and
// use_squared_matrix_, or the DDFT's dynamical matrix with the DDFT's use_squared_matrix_ (Dynamical_Matrix::
// matrix_dot_v, src/Dynamical_Matrix.cpp:34-50, is the DDFT's; its intern product uses the DFT's full_hessian_).
static int matrix_flags(bool full_hessian, bool dft_squared, const DDFT* ddft) {
  const bool squared = ddft ? ddft->get_use_squared_matrix_() : dft_squared;
  return (full_hessian ? 0 : CDFT_NO_IDEAL) | (squared ? CDFT_SQUARED_MATRIX : 0) | (ddft ? CDFT_DYNAMIC_MATRIX : 0) |
         (ddft && ddft->get_central_differences() ? CDFT_CENTRAL_DIFFERENCES : 0);
}

int Eigenvalues::flags() const {
  return matrix_flags(matrix_.full_hessian_, matrix_.use_squared_matrix_, dynamic_ ? ddft_ : nullptr);
}

void Eigenvalues::set_eigen_vec_(const std::vector<double>& v_in) {
  if (eigen_vec_.size() != (long)v_in.size()) eigen_vec_.zeros((long)v_in.size());
  for (long i = 0; i < (long)v_in.size(); ++i) eigen_vec_.set(i, v_in[i]);
}

double Eigenvalues::calculate_gradients(DFT_Vec& df) {
  const Species* s = matrix_.species_.front();
  const long N = s->getDensity().Ntot();
  if (df.size() != N) df.zeros(N);
  const int flags = this->flags();
  double f = 0.0;
  check(cdft_system_eigen_objective(matrix_.system(), s->getDensity().get_density_pointer(), eigen_vec_.memptr(),
                                    vshift_.size() == N ? vshift_.memptr() : nullptr, scale_, change_sign_ ? 1 : 0,
                                    flags, &f, df.memptr()));
  ++num_eval_;
  if (verbose_) std::printf("\r\tEvaluation %d gives estimated eigenvalue %g   ", num_eval_, (change_sign_ ? -1 : 1) * f / scale_);
  return f;
}

void Eigenvalues::calculate_eigenvector() {
  const Species* sp = matrix_.species_.front();
  const long N = sp->getDensity().Ntot();
  if (eigen_vec_.size() != N) {  // src/Eigenvalues.cpp:78-88
    eigen_vec_.zeros(N);
    std::mt19937_64 rng(seed_);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (long i = 0; i < N; ++i) eigen_vec_.set(i, nd(rng));
    if (sp->is_fixed_boundary()) {
      const Lattice& L = sp->getLattice();
      for (long i = 0; i < N; ++i) {
        const long ix = i / (L.Ny() * L.Nz()), iy = (i / L.Nz()) % L.Ny(), iz = i % L.Nz();
        if (ix == 0 || iy == 0 || iz == 0) eigen_vec_.set(i, 0.0);
      }
    }
  }
  // limited-memory BFGS (8 pairs) with a backtracking Armijo line search; stops when the relative change of f is
  // at most tol_ (nlopt's ftol_rel)
  const int m = 8;
  std::vector<DFT_Vec> S, Y;
  std::vector<double> rho_h;
  DFT_Vec x, g(N), xn, gn(N), d(N);
  x.set(eigen_vec_);
  double f = calculate_gradients(g);
  return_code_ = 5;
  while (num_eval_ < max_eval_) {
    d.set(g);
    d.MultBy(-1.0);
    std::vector<double> alpha(S.size());
    for (int i = (int)S.size() - 1; i >= 0; --i) {
      alpha[i] = rho_h[i] * S[i].dotWith(d);
      d.IncrementBy_Scaled_Vector(Y[i], -alpha[i]);
    }
    if (!S.empty()) d.MultBy(S.back().dotWith(Y.back()) / Y.back().dotWith(Y.back()));
    for (size_t i = 0; i < S.size(); ++i) {
      const double beta = rho_h[i] * Y[i].dotWith(d);
      d.IncrementBy_Scaled_Vector(S[i], alpha[i] - beta);
    }
    double gd = g.dotWith(d);
    if (!(gd < 0)) {  // not a descent direction: restart along -g
      S.clear();
      Y.clear();
      rho_h.clear();
      d.set(g);
      d.MultBy(-1.0);
      gd = g.dotWith(d);
    }
    double step = S.empty() ? 1.0 / std::max(1e-300, g.euclidean_norm()) * x.euclidean_norm() : 1.0;
    double fn = f;
    bool moved = false;
    for (int ls = 0; ls < 60 && num_eval_ < max_eval_; ++ls) {
      xn.set(x);
      xn.IncrementBy_Scaled_Vector(d, step);
      eigen_vec_.set(xn);
      fn = calculate_gradients(gn);
      if (fn <= f + 1e-4 * step * gd) {
        moved = true;
        break;
      }
      step *= 0.5;
    }
    if (!moved) break;
    DFT_Vec sv, yv;
    sv.set(xn);
    sv.IncrementBy_Scaled_Vector(x, -1.0);
    yv.set(gn);
    yv.IncrementBy_Scaled_Vector(g, -1.0);
    const double sy = sv.dotWith(yv);
    if (sy > 0) {
      if ((int)S.size() == m) {
        S.erase(S.begin());
        Y.erase(Y.begin());
        rho_h.erase(rho_h.begin());
      }
      S.push_back(sv);
      Y.push_back(yv);
      rho_h.push_back(1.0 / sy);
    }
    const double df_rel = std::fabs(fn - f);
    x.set(xn);
    g.set(gn);
    f = fn;
    if (df_rel <= tol_ * std::fabs(f)) {
      return_code_ = 3;
      break;
    }
  }
  eigen_vec_.set(x);
  eigen_val_ = f / scale_;
  if (change_sign_) eigen_val_ *= -1;
  eigen_vec_.normalise();
  if (verbose_) std::printf("\n");
}

double Log_Det::calculate_log_det(long seed, int num_samples, bool has_zero_eigenvalue, double& variance) {
  const Species* s = matrix_.species_.front();
  double ld = 0.0;
  const int flags = matrix_flags(matrix_.full_hessian_, matrix_.use_squared_matrix_, dynamic_ ? ddft_ : nullptr);
  check(cdft_system_log_det(matrix_.system(), s->getDensity().get_density_pointer(), order_, lam_max_, lam_min_, seed,
                            num_samples, has_zero_eigenvalue ? 1 : 0, flags, &ld, &variance));
  if (verbose_) std::printf("\t samples = %d log_det = %g variance = %g\n", num_samples, ld, variance);
  return ld;
}

double DFT::mu_times_beta(double density) const {
  double f, mu;
  check(cdft_system_bulk(system(), density, &f, &mu));
  return mu;
}
double DFT::fhelmholtz_times_beta_over_volume(double density) const {
  double f, mu;
  check(cdft_system_bulk(system(), density, &f, &mu));
  return f;
}
double DFT::omega_times_beta_over_volume(double density) const {
  return fhelmholtz_times_beta_over_volume(density) - density * mu_times_beta(density);
}

// Iterates the estimates of the critical point xc = 0.13044 (6/pi) d^-3, Tc = -0.090082 (2 a(Tc)) d^-3 Tc with
// d = hsd(Tc) (or HSD) until Tc changes by at most 1e-8 (T1 + Tc): DFT::getCriticalPoint (src/DFT_Coex.cpp:39-50).
void DFT::getCriticalPoint(Potential1& p, double& xc, double& Tc, double HSD) const {
  for (;;) {
    const double T1 = Tc;
    const double d = HSD < 0 ? p.getHSD(Tc) : HSD;
    xc = 0.13044 * (6.0 / M_PI) / (d * d * d);
    Tc = -0.090082 * 2 * p.getVDW_Parameter(Tc) / (d * d * d) * Tc;
    if (!(std::fabs(T1 - Tc) > 1e-8 * (T1 + Tc))) break;
  }
}

namespace {
// Golden-section search of [a, b] for a maximum (sense = 1) or minimum (sense = -1) of f, until b - a <= tol;
// returns the midpoint (the refinement steps of src/DFT_Coex.cpp:74-98, 113-131).
template <typename Fn>
double golden_extremum(Fn f, double a, double b, int sense, double tol) {
  const double r = (3 - std::sqrt(5.0)) / 2;
  double u = a + r * (b - a), v = b - r * (b - a);
  double fu = f(u), fv = f(v);
  while (b - a > tol) {
    if (sense * fu > sense * fv) {
      b = v;
      v = u;
      fv = fu;
      u = a + r * (b - a);
      fu = f(u);
    } else {
      a = u;
      u = v;
      fu = fv;
      v = b - r * (b - a);
      fv = f(v);
    }
  }
  return (a + b) / 2;
}
}  // namespace

// The stationary points of the pressure: its local maximum xs1 (refined to 1e-8) and the following local minimum
// xs2 (refined to tol), each bracketed by stepping dx: DFT::findSpinodal (src/DFT_Coex.cpp:53-134).
void DFT::findSpinodal(double xmax, double dx, double& xs1, double& xs2, double tol) const {
  auto P = [&](double x) { return -omega_times_beta_over_volume(x); };
  double x = 2 * dx, p = P(x);
  double dp = p - P(dx);
  if (dp < 0) throw std::runtime_error("0. Could not get started in DFT::findSpinodal");
  while (dp > 0 && x < xmax - dx) {
    const double q = P(x + dx);
    dp = q - p;
    p = q;
    x += dx;
  }
  if (x >= xmax - dx) throw std::runtime_error("1. Xmax exceeded in DFT::findSpinodal");
  double a = x - 2 * dx, b = x;
  double fs = P(a + dx);
  if (fs < P(a) || fs < P(b)) throw std::runtime_error("Failed sanity check 1 in DFT::findSpinodal");
  xs1 = golden_extremum(P, a, b, 1, 1e-8);
  x = xs1;
  p = P(x);
  do {
    const double q = P(x + dx);
    dp = q - p;
    p = q;
    x += dx;
  } while (dp < 0 && x < xmax - dx);
  if (x >= xmax - dx) throw std::runtime_error("2. Xmax exceeded in DFT::findSpinodal");
  a = x - 2 * dx;
  b = x;
  fs = P(a + dx);
  if (fs > P(a) || fs > P(b)) throw std::runtime_error("Failed sanity check 2 in DFT::findSpinodal");
  xs2 = golden_extremum(P, a, b, -1, tol);
}

// Bisection of mu(x) = mu on [xmin, xmax] until the bracketing mu values differ by at most tol:
// DFT::find_density_from_mu (src/DFT_Coex.cpp:136-157).
double DFT::find_density_from_mu(double mu, double xmin, double xmax, double tol) const {
  double mu1 = mu_times_beta(xmin), mu2 = mu_times_beta(xmax);
  if (mu1 > mu2) {
    std::swap(xmin, xmax);
    std::swap(mu1, mu2);
  }
  if (mu2 < mu || mu1 > mu) throw std::runtime_error("DFT::find_density_from_mu failed");
  do {
    const double x = (xmin + xmax) / 2, m = mu_times_beta(x);
    if (m > mu) {
      xmax = x;
      mu2 = m;
    } else {
      xmin = x;
      mu1 = m;
    }
  } while (mu2 - mu1 > tol);
  return (xmin + xmax) / 2;
}

// Vapour-liquid coexistence: equal mu by find_density_from_mu on the liquid branch [xs2, xmax - dx], equal pressure
// by shrinking the vapour density by 1.1 until the pressure difference changes sign, then bisection to tol:
// DFT::findCoex (src/DFT_Coex.cpp:159-196).
void DFT::findCoex(double xmax, double dx, double& x1, double& x2, double tol) const {
  double xs1, xs2;
  findSpinodal(xmax, dx, xs1, xs2, tol);
  auto P = [&](double x) { return -omega_times_beta_over_volume(x); };
  auto liquid = [&](double xv) { return find_density_from_mu(mu_times_beta(xv), xs2, xmax - dx, tol); };
  double xvap = xs1, xliq = liquid(xvap);
  double dp1 = P(xvap) - P(xliq), dp2;
  do {
    dp2 = dp1;
    xvap /= 1.1;
    xliq = liquid(xvap);
    dp1 = P(xvap) - P(xliq);
  } while (xvap > 1e-16 && ((dp1 < 0) == (dp2 < 0)));
  if (xvap < 1e-16) throw std::runtime_error("DFT::findCoex failed 1");
  double y1 = xvap, y2 = xvap * 1.1;
  do {
    const double y = (y1 + y2) / 2;
    xliq = liquid(y);
    const double dp = P(y) - P(xliq);
    if ((dp < 0) == (dp1 < 0)) {
      y1 = y;
      dp1 = dp;
    } else {
      y2 = y;
    }
  } while (std::fabs(y2 - y1) > tol);
  x1 = (y1 + y2) / 2;
  x2 = xliq;
}

// ---------------------------------------------------------------- Minimizer
// The species takes the minimizer's density and, with a fixed mass, the multiplier of its last evaluation.
void Minimizer::sync_species() {
  const int S = dft_->getNumberOfSpecies();
  const long N = dft_->get_lattice().Ntot();
  DFT_Vec rho(N * S);
  density(rho);
  if (S > 1) {  // species after species, with the fixed-mass multipliers of the last evaluation
    std::vector<double> mu(S);
    check(cdft_mixture_chem_potential(dft_->mixture(), mu.data()));
    for (int i = 0; i < S; ++i) {
      Species* s = dft_->getSpecies(i);
      for (long j = 0; j < N; ++j) s->set_density(j, rho.get(i * N + j));
      s->endForceCalculation(mu[i]);
    }
    return;
  }
  Species* s = dft_->getSpecies(0);
  s->set_density(rho);
  double mu = 0.0;
  check(cdft_system_chem_potential(dft_->system(), &mu));
  s->endForceCalculation(mu);
}

// Minimizer::run / resume, src/Minimizer.cpp:36-87 (no wall-clock or snapshot I/O)
void Minimizer::run(long maxSteps) {
  step_counter_ = 0;
  calls_ = 0;
  resume(maxSteps);
}

void Minimizer::resume(long maxSteps) {
  do {
    F_ = step();
    step_counter_++;
    const Density& D = dft_->getSpecies(0)->getDensity();
    const double Ntotal = D.getNumberAtoms(), Volume = D.getVolume();
    if (get_convergence_monitor() < forceLimit_) break;
    if (maxSteps > 0 && step_counter_ == maxSteps) break;
    if (Ntotal / Volume < minDensity_) break;
  } while (true);
}

void Minimizer::set_fixed_direction(const DFT_Vec& fixed) { fixed_direction_.set(fixed); }

fireMinimizer2::fireMinimizer2(DFT* dft) : Minimizer(dft) {
  cdft_fire_default_params(&params_);
  forceLimit_ = 0.1;
}
fireMinimizer2::~fireMinimizer2() {
  if (fire_) cdft_fire_destroy(fire_);
}
// The species' densities one after the other (FIRE over several species runs on the concatenated variables).
static DFT_Vec all_densities(const DFT& dft) {
  const int S = dft.getNumberOfSpecies();
  const long N = dft.get_lattice().Ntot();
  DFT_Vec r(N * S);
  for (int i = 0; i < S; ++i) std::memcpy(r.memptr() + i * N, dft.getDensity(i).get_density_pointer(), sizeof(double) * N);
  return r;
}

void fireMinimizer2::ensure() {
  if (fire_) return;
  const int S = dft_->getNumberOfSpecies();
  const long Nall = dft_->get_lattice().Ntot() * S;
  if (S > 1) {
    std::vector<const double*> r;
    for (int i = 0; i < S; ++i) r.push_back(dft_->getDensity(i).get_density_pointer());
    check(cdft_mixture_fire_create(dft_->mixture(), r.data(), 0, &params_, &fire_));
  } else {
    check(cdft_fire_create(dft_->system(), dft_->getDensity(0).get_density_pointer(), 0, &params_, &fire_));
  }
  if (fixed_direction_.size() == Nall) {
    // the constructor's getDF_DX already reflects the force: restart with the direction in place
    check(cdft_fire_set_fixed_direction(fire_, fixed_direction_.memptr(), 0));
    check(cdft_fire_reset(fire_, all_densities(*dft_).memptr(), 0));
  }
}
void fireMinimizer2::reset() {
  if (!fire_) return ensure();
  const long Nall = dft_->get_lattice().Ntot() * dft_->getNumberOfSpecies();
  check(cdft_fire_set_fixed_direction(fire_, fixed_direction_.size() == Nall ? fixed_direction_.memptr() : nullptr,
                                      0));
  const int rc = cdft_fire_reset(fire_, all_densities(*dft_).memptr(), 0);
  if (rc < 0) check(rc);
  if (rc == CDFT_ETA_TOO_LARGE) throw Eta_Too_Large_Exception();
  sync_species();
}
double fireMinimizer2::step() {
  ensure();
  const int rc = cdft_fire_step(fire_, &rep_);
  if (rc == CDFT_UPHILL_LIMIT) throw std::runtime_error("Cannot stop going uphill in Fire2");
  if (rc < 0) check(rc);
  calls_ = rep_.calls;
  F_ = rep_.F;
  sync_species();
  if (snapshot_every_ > 0 && snapshot_count_ % snapshot_every_ == 0) {
    std::string of = "snapshot_s0.dat";
    dft_->writeDensity(0, of);
  }
  ++snapshot_count_;
  return F_;
}
int fireMinimizer2::run_to_raw_tolerance(long maxSteps, double tol) {
  ensure();
  const int rc = cdft_fire_run(fire_, (int)maxSteps, tol, &rep_);
  if (rc < 0) check(rc);
  F_ = rep_.F;
  calls_ = rep_.calls;
  sync_species();
  return rep_.steps;
}
void fireMinimizer2::density(DFT_Vec& out) const {
  out.resize(dft_->get_lattice().Ntot() * dft_->getNumberOfSpecies());
  check(cdft_fire_get(fire_, out.memptr(), nullptr, nullptr, nullptr));
}

DDFT::DDFT(DFT* dft, bool, bool central_differences) : Minimizer(dft), central_(central_differences) {
  const double dx = dft->get_lattice().getDX();
  dt_ = 0.0001 * dx * dx;  // src/DDFT.cpp:46
}
DDFT::~DDFT() {
  if (ddft_) cdft_ddft_destroy(ddft_);
}
void DDFT::ensure() {
  if (ddft_) return;
  check(cdft_ddft_create(dft_->system(), dft_->getSpecies(0)->getDensity().get_density_pointer(), dt_, dtMax_, tol_,
                         &ddft_));
  if (central_) check(cdft_ddft_set_central_differences(ddft_, 1));
}
void DDFT::set_max_time_step(double t) {
  dtMax_ = t;
  if (ddft_) check(cdft_ddft_set_max_time_step(ddft_, t));
}
double DDFT::step() {
  ensure();
  check(cdft_ddft_step(ddft_, &rep_));
  F_ = rep_.F;
  calls_ = rep_.calls;
  sync_species();
  return F_;
}
DDFT_IF_Open::DDFT_IF_Open(DFT* dft, double background, bool) : Minimizer(dft), background_(background) {
  const double dx = dft->get_lattice().getDX();
  dt_ = 0.0001 * dx * dx;
}
DDFT_IF_Open::~DDFT_IF_Open() {
  if (ddft_) cdft_ddft_open_destroy(ddft_);
}
void DDFT_IF_Open::ensure() {
  if (ddft_) return;
  check(cdft_ddft_open_create(dft_->system(), dft_->getSpecies(0)->getDensity().get_density_pointer(), background_,
                              dt_, dtMax_, tol_, &ddft_));
}
double DDFT_IF_Open::step() {
  ensure();
  check(cdft_ddft_open_step(ddft_, &rep_));
  F_ = rep_.F;
  calls_ = rep_.calls;
  sync_species();
  return F_;
}
void DDFT_IF_Open::density(DFT_Vec& out) const {
  out.resize(dft_->get_lattice().Ntot());
  check(cdft_ddft_open_get_density(ddft_, out.memptr()));
}

void DDFT::density(DFT_Vec& out) const {
  out.resize(dft_->get_lattice().Ntot());
  check(cdft_ddft_get_density(ddft_, out.memptr()));
}
