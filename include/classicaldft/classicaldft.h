// classicaldft.h - source-level drop-in for the hot path of classicalDFT.
//
// Classes carry the reference's names and the signatures a caller of the hot
// path uses (Lattice, Density, Species, FMT_Species, FMT and its five models,
// Potential1/LJ, Interaction_Interpolation_QF, DFT, Minimizer/fireMinimizer2,
// DDFT, Eta_Too_Large_Exception).  Every evaluation and every solver step runs
// on the GPU through the C ABI of <cdft.h>; these classes only hold host
// copies of the fields the reference exposes and translate status codes into
// the reference's exceptions.  References are to the classicalDFT sources.
#ifndef CLASSICALDFT_AMD_H
#define CLASSICALDFT_AMD_H

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../cdft.h"

// --------------------------------------------------------------- vectors ---
// DFT_Vec (include/DFT_LinAlg.h:14-97): the operations the hot path's callers use.
class DFT_Vec {
 public:
  DFT_Vec() = default;
  explicit DFT_Vec(long n) : d_(n, 0.0) {}
  long size() const { return (long)d_.size(); }
  void resize(long n) { d_.resize(n); }
  void zeros() { std::fill(d_.begin(), d_.end(), 0.0); }
  void zeros(long n) { d_.assign(n, 0.0); }
  double get(long i) const { return d_[i]; }
  void set(long i, double v) { d_[i] = v; }
  void set(const DFT_Vec& o) { d_ = o.d_; }
  void IncrementBy(long i, double v) { d_[i] += v; }
  void IncrementBy(const DFT_Vec& o);
  void IncrementBy_Scaled_Vector(const DFT_Vec& o, double s);
  void MultBy(double s);
  void add(double s);
  double inf_norm() const;
  double euclidean_norm() const;
  double dotWith(const DFT_Vec& o) const;
  double accu() const;
  double min() const;
  double max() const;
  double* memptr() { return d_.data(); }
  const double* memptr() const { return d_.data(); }
  void normalise() { MultBy(1.0 / euclidean_norm()); }

 private:
  std::vector<double> d_;
};

// ------------------------------------------------------------ exceptions ---
// include/FMT.h:649-663
class Eta_Too_Large_Exception : public std::runtime_error {
 public:
  Eta_Too_Large_Exception() : std::runtime_error("Eta too large in FMT::dPHI") {}
};

// --------------------------------------------------------------- lattice ---
// include/Lattice.h:11-134
class Lattice {
 public:
  Lattice(double dx, double L[]) : Lattice(dx, dx, dx, L) {}
  Lattice(double dx, double dy, double dz, double L[]);
  double getDX() const { return dx_; }
  double getDY() const { return dy_; }
  double getDZ() const { return dz_; }
  long Nx() const { return Nx_; }
  long Ny() const { return Ny_; }
  long Nz() const { return Nz_; }
  long Ntot() const { return Nx_ * Ny_ * Nz_; }
  long size() const { return Ntot(); }
  long Nout() const { return Nx_ * Ny_ * (Nz_ / 2 + 1); }
  double dV() const { return dx_ * dy_ * dz_; }
  double getVolume() const { return L_[0] * L_[1] * L_[2]; }
  double Lx() const { return L_[0]; }
  double Ly() const { return L_[1]; }
  double Lz() const { return L_[2]; }
  long pos(int i, int j, int k) const;
  long get_PBC_Pos(int ix, int iy, int iz) const;

 protected:
  long Nx_ = 0, Ny_ = 0, Nz_ = 0;
  double dx_ = 0, dy_ = 0, dz_ = 0;
  double L_[3] = {0, 0, 0};
};

// --------------------------------------------------------------- density ---
// include/Density.h:35-118; the external field gets a public setter.
class Density : public Lattice {
 public:
  Density(double dx, double L[]);
  Density(double dx, double dy, double dz, double L[]);
  void set(long pos, double val) { rho_.set(pos, val); }
  void set(int i, int j, int k, double val) { rho_.set(pos(i, j, k), val); }
  void set(const DFT_Vec& x) { rho_.set(x); }
  void set(double val);
  double get(long pos) const { return rho_.get(pos); }
  double getNumberAtoms() const;  // compensated, src/Density.cpp:136-142
  void scale_to(double a) { rho_.MultBy(a); }
  const DFT_Vec& get_density_real() const { return rho_; }
  const double* get_density_pointer() const { return rho_.memptr(); }
  const DFT_Vec& get_external_field() const { return vWall_; }
  void set_external_field(const DFT_Vec& v) { vWall_ = v; ++field_version_; }
  bool has_external_field() const { return has_field_(); }
  unsigned field_version() const { return field_version_; }
  // src/Density.cpp:227-244 with DFT_Vec::save / load (additionalCode/DFT_Eigen.cpp:71-85): text, "N" on the first
  // line, then the N values separated by blanks at the stream's default precision (6 significant digits).
  // readDensity throws std::runtime_error if the file cannot be opened or holds another number of values.
  void writeDensity(std::string filename) const;
  void readDensity(const char* filename);

 private:
  bool has_field_() const;
  DFT_Vec rho_;
  DFT_Vec vWall_;
  unsigned field_version_ = 0;
};

// --------------------------------------------------------------- species ---
// include/Species.h:13-97, src/Species.cpp
class Species {
 public:
  Species(Density& density, double mu = 0, int seq = -1);
  virtual ~Species() = default;
  int getSequenceNumber() const { return seq_num_; }  // include/Species.h:16-19
  const Lattice& getLattice() const { return density_; }
  const Density& getDensity() const { return density_; }
  Density& density() { return density_; }
  void set_density(DFT_Vec& x) { density_.set(x); }
  void set_density(long j, double x) { density_.set(j, x); }
  double getChemPotential() const { return mu_; }
  void setChemPotential(double m) { mu_ = m; fixedMass_ = -1; }
  void setFixedMass(double m) { fixedMass_ = m; if (m > 0.0) mu_ = 0.0; }
  bool is_mass_fixed() const { return fixedMass_ > 0; }
  double getFixedMass() const { return fixedMass_; }
  // include/Species.h:24-29: forces zeroed on (fixed background) or averaged over (homogeneous boundary) the
  // boundary points ix, iy or iz = 0; applied on the GPU by the full evaluation
  void setFixedBackground(bool fixed) { fixedBackground_ = fixed; }
  void setHomogeneousBoundary(bool homo) { homogeneousBoundary_ = homo; }
  bool is_background_fixed() const { return fixedBackground_; }
  bool is_fixed_boundary() const { return fixedBackground_; }
  bool is_boundary_homogeneous() const { return homogeneousBoundary_; }
  void zeroForce() { dF_.zeros(); }
  void addToForce(const DFT_Vec& f) { dF_.IncrementBy(f); }
  void addToForce(long i, double v) { dF_.IncrementBy(i, v); }
  DFT_Vec& getDF() { return dF_; }
  const DFT_Vec& get_const_DF() const { return dF_; }
  void setDF(DFT_Vec& df) { dF_.set(df); }
  double get_convergence_monitor() const { return dF_.inf_norm() / density_.dV(); }
  virtual double getHSD() const { return 0.0; }
  void beginForceCalculation();   // src/Species.cpp:64-71
  // src/Species.cpp:73-112: the boundary conditions and the fixed-mass projection of dF run on the GPU inside
  // the evaluation; this records the fixed-mass multiplier it returns as mu (Species::getChemPotential).
  double endForceCalculation(double multiplier);

 protected:
  Density& density_;
  DFT_Vec dF_;
  double mu_;
  double fixedMass_ = -1;
  bool fixedBackground_ = false, homogeneousBoundary_ = false;
  int seq_num_ = 0;
  static int SequenceNumber_;
};

// include/Species.h:109-373: the weights live on the GPU (resident spectra).
class FMT_Species : public Species {
 public:
  FMT_Species(Density& density, double hsd, double mu = 0, int seq = -1);
  double getHSD() const override { return hsd_; }
  // Asakura-Oosawa polymer parameters (0 for a plain FMT species)
  virtual double ao_Rp() const { return 0.0; }
  virtual double ao_reservoir() const { return 0.0; }
  virtual int ao_force() const { return CDFT_AO_FORCE_EXACT; }

 private:
  double hsd_;
};

// include/Species.h:380-461, src/FMT_Species.cpp:647-738: an FMT species with ideal polymer of radius Rp at
// reservoir density reservoir_density (FP64).  The AO force counts the tensor off-diagonals twice, as the
// derivative of the AO free energy requires (the reference stores one of the two entries).
class FMT_AO_Species : public FMT_Species {
 public:
  FMT_AO_Species(Density& density, double hsd, double Rp, double reservoir_density, double mu = 0, int seq = -1)
      : FMT_Species(density, hsd, mu, seq), Rp_(Rp), reservoir_density_(reservoir_density) {}
  double ao_Rp() const override { return Rp_; }
  double ao_reservoir() const override { return reservoir_density_; }
  int ao_force() const override { return reference_force_ ? CDFT_AO_FORCE_REFERENCE : CDFT_AO_FORCE_EXACT; }
  // true: the force counts each off-diagonal tensor measure once, as setFundamentalMeasures_AO does
  // (include/Species.h:406-425); false (default): the derivative of the AO free energy.
  void setReferenceForce(bool on) { reference_force_ = on; }
  double getReservoirDensity() const { return reservoir_density_; }
  double getHSDP() const { return 2 * Rp_; }

 private:
  double Rp_, reservoir_density_;
  bool reference_force_ = false;
};

// ------------------------------------------------------------------- FMT ---
// include/FMT.h:28-643.  calculateFreeEnergyAndDerivatives accumulates into
// each FMT species' dF and returns F_hs (x dV), like src/FMT.cpp:363-427.
class FMT {
 public:
  virtual ~FMT();
  double calculateFreeEnergyAndDerivatives(std::vector<Species*>& allSpecies);
  // src/FMT.cpp:437-567: d2F[s] += (d2 F_hs / drho drho) v[s] (directions in real space; the reference takes
  // their DFT_FFT with the spectrum already computed).
  void add_second_derivative(const std::vector<DFT_Vec>& v, std::vector<DFT_Vec>& d2F,
                             const std::vector<Species*>& allSpecies);
  double BulkMuex(const std::vector<double>& x, const std::vector<Species*>& allSpecies, int species) const;
  double BulkFex(const std::vector<double>& x, const std::vector<Species*>& allSpecies) const;
  virtual std::string Name() const = 0;
  std::string get_name() const { return Name(); }
  virtual bool needsTensor() const { return model_ >= CDFT_ESFMT; }
  int model_id() const { return model_; }
  double A() const { return A_; }
  double B() const { return B_; }
  void set_precision(int p) { precision_ = p; }  // CDFT_FP64 (default) or CDFT_MIXED
  int precision() const { return precision_; }

 protected:
  FMT(int model, double A = 1, double B = 0) : model_(model), A_(A), B_(B) {}
  int model_;
  double A_, B_;
  int precision_ = CDFT_FP64;
  mutable cdft_system sys_ = nullptr;
  mutable const Lattice* sys_lattice_ = nullptr;
  mutable double sys_hsd_ = -1;
  mutable int sys_aof_ = -1;
  mutable const Species* sys_species_ = nullptr;
  cdft_system system_for(const Species& s) const;
  // several FMT species: a hard-sphere mixture (cdft_mixture), rebuilt when the species or their diameters change
  mutable cdft_mixture mix_ = nullptr;
  mutable std::vector<const FMT_Species*> mix_species_;
  mutable std::vector<double> mix_hsd_;
  cdft_mixture mixture_for(const std::vector<FMT_Species*>& species) const;
};

class Rosenfeld : public FMT {
 public:
  Rosenfeld() : FMT(CDFT_ROSENFELD) {}
  std::string Name() const override { return "Rosenfeld"; }
};
class RSLT : public FMT {
 public:
  RSLT() : FMT(CDFT_RSLT) {}
  std::string Name() const override { return "RSLT"; }
};
class esFMT : public FMT {
 public:
  esFMT(double A = 1, double B = 0) : FMT(CDFT_ESFMT, A, B) {}
  std::string Name() const override;
};
class WhiteBearI : public FMT {
 public:
  WhiteBearI() : FMT(CDFT_WHITEBEAR_I, 1, -1) {}
  std::string Name() const override { return "WhiteBear I"; }
};
class WhiteBearII : public FMT {
 public:
  WhiteBearII() : FMT(CDFT_WHITEBEAR_II, 1, -1) {}
  std::string Name() const override { return "WhiteBear II"; }
};

// ------------------------------------------------------------ potentials ---
// include/Potential1.h:23-189 (LJ with the WCA split; BH split not provided)
class Potential1 {
 public:
  Potential1(double sigma, double eps, double rcut) : sigma_(sigma), eps_(eps), rcut_(rcut) {}
  virtual ~Potential1() = default;
  double getRcut() const { return rcut_; }
  double getSigma() const { return sigma_; }
  double getEps() const { return eps_; }  // the constructor's eps (WHDF rescales it internally)
  double getR0() const { return r0_; }
  virtual double getRmin() const = 0;
  virtual double getHardCore() const { return 0.0; }
  void set_WCA_limit(double r) { r_att_min_ = r; }           // Potential1.h:46
  void setBH() { bh_ = true; r_att_min_ = getR0(); }         // Potential1.h:47
  double V(double r) const { return vr(r) - shift_; }
  double Watt(double r) const;                               // Potential1.h:49, 67-75
  double V0(double r) const;                                 // Potential1.h:77-81
  // Potential1.h:73-91: adaptive Gauss-Kronrod to 1e-13 instead of GSL qags at 1e-4 / 1e-6
  double getHSD(double kT) const;
  double getVDW_Parameter(double kT) const;
  // the C ABI description of this potential and split
  cdft_potential_desc desc() const;
  virtual int kind() const = 0;
  virtual double alpha() const { return 0.0; }
  // include/Potential1.h:160-165, 227-232, 297-302 (rcut and the WCA limit printed without a separator)
  std::string getIdentifier() const;
  bool isBH() const { return bh_; }
  double getWCALimit() const { return r_att_min_; }

 protected:
  virtual double vr(double r) const = 0;
  double sigma_, eps_, rcut_;
  double shift_ = 0, rmin_ = 1, Vmin_ = -1, r0_ = 1, r_att_min_ = 0;
  bool bh_ = false;
};

// include/Potential1.h:136-189
class LJ : public Potential1 {
 public:
  LJ(double sigma, double eps, double rcut);
  double getRmin() const override { return std::pow(2.0, 1.0 / 6.0) * sigma_; }
  int kind() const override { return CDFT_POT_LJ; }

 protected:
  double vr(double r) const override {
    const double y = sigma_ / r, y3 = y * y * y, y6 = y3 * y3;
    return 4 * eps_ * (y6 * y6 - y6);
  }
};

// include/Potential1.h:195-257
class tWF : public Potential1 {
 public:
  tWF(double sigma, double eps, double rcut, double alpha = 50.0);
  double getRmin() const override { return sigma_ * std::sqrt(1 + std::pow(2.0 / alpha_, 1.0 / 3.0)); }
  double getHardCore() const override { return sigma_; }
  int kind() const override { return CDFT_POT_TWF; }
  double alpha() const override { return alpha_; }

 protected:
  double vr(double r) const override;
  double alpha_;
};

// include/Potential1.h:262-324
class WHDF : public Potential1 {
 public:
  WHDF(double sigma, double eps, double rcut);
  double getRmin() const override { return rcut_ * std::pow((1.0 + 2.0 * (rcut_ / sigma_) * (rcut_ / sigma_)) / 3.0, -0.5); }
  int kind() const override { return CDFT_POT_WHDF; }

 protected:
  double vr(double r) const override;
  double eps_eff_;
};

// ----------------------------------------------------------- interaction ---
// include/Interaction.h:33-431: lattice mean-field kernels (single species, s1 == s2).
class Interaction_Base {
 public:
  Interaction_Base(Species* s1, Species* s2, Potential1* v, double kT, int scheme = CDFT_MF_QF, int ngauss = 0);
  virtual ~Interaction_Base() = default;
  double getVDWParameter() const;  // a_vdw / kT of the lattice kernel
  double Mu(const std::vector<double>& x, int species) const;
  double Fhelmholtz(const std::vector<double>& x) const;
  Species* species1() const { return s1_; }
  Potential1* potential() const { return v_; }
  double kT() const { return kT_; }
  // the mean-field fields of a system description (potential, split, rule, kT)
  void describe(cdft_system_desc& d) const;
  // src/Interaction.cpp:54-59: multiplies the kernel and a_vdw (DFTs built afterwards use the scaled kernel)
  void scale_interaction(double scale_fac) { scale_ *= scale_fac; a_vdw_ = NAN; w_.clear(); }
  double scale() const { return scale_; }
  // include/Interaction.h:41: the kernel value (already / kT, scaled) at lattice position s
  double getW(long s) const;
  // include/Interaction.h:57
  std::string get_name() const { return v_ ? v_->getIdentifier() : std::string("none"); }

 protected:
  virtual void ensure_kernel() const {}
  Species *s1_, *s2_;
  Potential1* v_;
  double kT_;
  int scheme_, ngauss_;
  mutable double a_vdw_ = NAN;
  mutable std::vector<double> kernel_;  // explicit kernel (already / kT) of the points-file class
  mutable std::vector<double> w_;       // host copy for getW
  double scale_ = 1.0;
};

// include/Interaction.h:97-123, src/Interaction.cpp:566-860: the kernel from a file of spherical integration
// points (generated on the GPU), cached in weights_<seq1>_<seq2>_<Nx>_<Ny>_<Nz>_<identifier>_.dat in the working
// directory in the reference's format and read back when the grid, spacing, potential and points file match.
class Interaction : public Interaction_Base {
 public:
  Interaction(Species* s1, Species* s2, Potential1* v, double kT, std::string pointsFile);
  bool readWeights();       // src/Interaction.cpp:589-656
  void generateWeights();   // src/Interaction.cpp:672-860, writes the cache
  bool weights_from_cache() const { return from_cache_; }
  const std::string& weights_file() const { return weightsFile_; }

 protected:
  void ensure_kernel() const override;
  std::string pointsFile_, weightsFile_;
  bool from_cache_ = false;
};

class Interaction_Interpolation_QF : public Interaction_Base {  // Interaction.h:416-431
 public:
  Interaction_Interpolation_QF(Species* s1, Species* s2, Potential1* v, double kT)
      : Interaction_Base(s1, s2, v, kT, CDFT_MF_QF) {}
};
class Interaction_Interpolation_Zero : public Interaction_Base {  // Interaction.h:293-313
 public:
  Interaction_Interpolation_Zero(Species* s1, Species* s2, Potential1* v, double kT)
      : Interaction_Base(s1, s2, v, kT, CDFT_MF_ZERO) {}
};
class Interaction_Interpolation_LE : public Interaction_Base {  // Interaction.h:318-343
 public:
  Interaction_Interpolation_LE(Species* s1, Species* s2, Potential1* v, double kT)
      : Interaction_Base(s1, s2, v, kT, CDFT_MF_LE) {}
};
class Interaction_Interpolation_QE : public Interaction_Base {  // Interaction.h:348-373
 public:
  Interaction_Interpolation_QE(Species* s1, Species* s2, Potential1* v, double kT)
      : Interaction_Base(s1, s2, v, kT, CDFT_MF_QE) {}
};
class Interaction_Interpolation_LF : public Interaction_Base {  // Interaction.h:378-403
 public:
  Interaction_Interpolation_LF(Species* s1, Species* s2, Potential1* v, double kT)
      : Interaction_Base(s1, s2, v, kT, CDFT_MF_LF) {}
};
class Interaction_Gauss_E : public Interaction_Base {  // Interaction.h:168-201, src/Interaction.cpp:482-520
 public:
  Interaction_Gauss_E(Species* s1, Species* s2, Potential1* v, double kT, int Ngauss)
      : Interaction_Base(s1, s2, v, kT, CDFT_MF_GAUSS_E, Ngauss) {}
};
class Interaction_Gauss_F : public Interaction_Base {  // Interaction.h:203-236, src/Interaction.cpp:465-480
 public:
  Interaction_Gauss_F(Species* s1, Species* s2, Potential1* v, double kT, int Ngauss)
      : Interaction_Base(s1, s2, v, kT, CDFT_MF_GAUSS_F, Ngauss) {}
};

// ------------------------------------------------------------------- DFT ---
// include/DFT.h:35-160, src/DFT.cpp:58-425 (single species).
class DFT {
 public:
  explicit DFT(Species* s = nullptr);
  ~DFT();
  void addSpecies(Species* s);
  void addInteraction(Interaction_Base* i);
  void addHardCoreContribution(FMT* fmt);
  const Lattice& get_lattice() const { return species_.front()->getLattice(); }
  const Density& getDensity(int s) const { return species_[s]->getDensity(); }
  Species* getSpecies(int s) const { return species_[s]; }
  int getNumberOfSpecies() const { return (int)species_.size(); }
  double getNumberAtoms(int s) const { return species_[s]->getDensity().getNumberAtoms(); }
  DFT_Vec& getDF(int s) { return species_[s]->getDF(); }
  double calculateFreeEnergyAndDerivatives(bool onlyFex);
  void writeDensity(int i, std::string& of) const { species_[i]->getDensity().writeDensity(of); }  // DFT.h:76
  // Dynamical_Matrix interface (include/Dynamical_Matrix.h:22-25, src/Dynamical_Matrix.cpp:34-70,
  // src/DFT.cpp:435-488): result = (d2 beta Omega / drho drho) v, the ideal-gas term only with
  // full_hessian_; with use_squared_matrix_ the matrix is applied twice.  Periodic systems.
  void matrix_dot_v_intern(const std::vector<DFT_Vec>& v, std::vector<DFT_Vec>& result, void* param = nullptr) const;
  void matrix_dot_v(const std::vector<DFT_Vec>& v, std::vector<DFT_Vec>& result, void* param = nullptr) const;
  void matrix_dot_v1(const DFT_Vec& v, DFT_Vec& result, void* param = nullptr) const;
  void set_full_hessian(bool full) { full_hessian_ = full; }
  void set_use_squared_matrix_(bool v) { use_squared_matrix_ = v; }
  double mu_times_beta(double density) const;
  double fhelmholtz_times_beta_over_volume(double density) const;
  double omega_times_beta_over_volume(double density) const;
  // Bulk phase diagram of this functional (src/DFT_Coex.cpp:39-196), on cdft_system_bulk: pressure p = -omega,
  // chemical potential mu(x).
  void getCriticalPoint(Potential1& p, double& xc, double& Tc, double HSD = -1) const;
  void findSpinodal(double xmax, double dx, double& xs1, double& xs2, double tol) const;
  double find_density_from_mu(double mu, double xmin, double xmax, double tol) const;
  void findCoex(double xmax, double dx, double& x1, double& x2, double tol) const;
  double get_f_id() const { return F_id_; }
  double get_f_ext() const { return F_ext_; }
  double get_f_hs() const { return F_hs_; }
  double get_f_mf() const { return F_mf_; }
  // The GPU system holding this functional (built on first use; one species).
  cdft_system system() const;
  // Several species: the hard-sphere mixture holding this functional (cdft_mixture, built on first use): each
  // species' ideal gas, external field, mu and constraints (fixed mass, fixed boundary) beside the mixture's FMT
  // (src/DFT.cpp:360-425).  FMT species only, with no mean-field interaction (the reference's between-species term
  // throws, src/Interaction.cpp:110).
  cdft_mixture mixture() const;
  void set_precision(int p);

 private:
  std::vector<Species*> species_;
  std::vector<Interaction_Base*> interactions_;
  FMT* fmt_ = nullptr;
  int precision_ = CDFT_FP64;
  mutable cdft_system sys_ = nullptr;
  mutable double sys_mu_ = NAN, sys_mass_ = NAN, sys_mfs_ = NAN;
  mutable int sys_bnd_ = -1, sys_aof_ = -1;
  mutable unsigned sys_field_ = ~0u;
  mutable cdft_mixture mix_ = nullptr;
  mutable std::vector<double> mix_key_;  // diameters, mu and field versions the mixture was built with
  double F_id_ = 0, F_ext_ = 0, F_hs_ = 0, F_mf_ = 0;
  bool full_hessian_ = true, use_squared_matrix_ = false;
  void invalidate() const;
  friend class Log_Det;
  friend class Eigenvalues;
  friend class DDFT;
};

class DDFT;

// include/Eigenvalues.h:16-59, src/Eigenvalues.cpp: the smallest eigenvalue of the DFT's Hessian (the largest with
// set_change_sign) by minimising the scaled Rayleigh quotient, whose value and gradient come from the GPU
// (cdft_system_eigen_objective).  The reference minimises with nlopt's LD_LBFGS (absent here); this uses a
// limited-memory BFGS of its own with the same relative tolerance on f: the iterates differ, the converged eigenpair
// agrees.  get_return_code: 3 = converged on the f tolerance, 5 = evaluation limit (nlopt's FTOL_REACHED /
// MAXEVAL_REACHED).  calculate_eigenvector starts from get_eigen_vec() when it has Ntot entries, else from a
// seeded normal vector (set_seed) that is zero on a fixed boundary.
class Eigenvalues {
 public:
  Eigenvalues(const DFT& matrix, bool verbose = false) : matrix_(matrix), verbose_(verbose) {}
  // the DDFT's dynamical matrix g . H (DDFT::matrix_dot_v_intern, src/DDFT.cpp:353-361) instead of the Hessian
  explicit Eigenvalues(const DDFT& matrix, bool verbose = false);
  void set_scale(double d) { scale_ = d; }
  void set_tolerence(double d) { tol_ = d; }
  void set_change_sign(bool b) { change_sign_ = b; }
  void set_vshift(DFT_Vec& v) { vshift_ = v; }
  void set_seed(unsigned long s) { seed_ = s; }
  void set_max_evaluations(int n) { max_eval_ = n; }
  double get_scale() const { return scale_; }
  double get_tolerence() const { return tol_; }
  double get_eigenvalue() const { return eigen_val_; }
  bool get_change_sign() const { return change_sign_; }
  int get_num_eval() const { return num_eval_; }
  int get_return_code() const { return return_code_; }
  const DFT_Vec& get_eigen_vec() const { return eigen_vec_; }
  double get_eigen_val() const { return eigen_val_; }
  void calculate_eigenvector();
  void set_eigen_vec_(const std::vector<double>& v_in);
  double calculate_gradients(DFT_Vec& df);
  void clear() { eigen_vec_.zeros(1); }

 private:
  const DFT& matrix_;
  DFT_Vec vshift_, eigen_vec_;
  double scale_ = 1e6, tol_ = 1e-6, eigen_val_ = 0;
  bool change_sign_ = false, verbose_ = false;
  int num_eval_ = 0, return_code_ = 0, max_eval_ = 100000;
  unsigned long seed_ = 1;
  bool dynamic_ = false;
  const DDFT* ddft_ = nullptr;  // the DDFT whose matrix this is (dynamic_), for its use_squared_matrix_
  int flags() const;
};

// include/Log_Det.h:22-51, src/Log_Det.cpp:22-166: stochastic Chebyshev estimate of log det of the DFT's
// Hessian (its full_hessian_ and use_squared_matrix_ settings apply), every product on the GPU.  Built on a DDFT,
// the matrix is the DDFT's (its use_squared_matrix_, the DFT's full_hessian_).
class Log_Det {
 public:
  Log_Det(const DFT& matrix, int order, double lam_max, double lam_min, bool verbose = false)
      : matrix_(matrix), order_(order), lam_max_(lam_max), lam_min_(lam_min), verbose_(verbose) {}
  // the DDFT's dynamical matrix, each product negated (is_dynamic, src/Log_Det.cpp:124, 139)
  Log_Det(const DDFT& matrix, int order, double lam_max, double lam_min, bool verbose = false);
  double calculate_log_det(long seed, int num_samples, bool has_zero_eigenvalue, double& variance);
  void set_verbose(bool b) { verbose_ = b; }
  void set_debug(bool b) { debug_ = b; }

 private:
  const DFT& matrix_;
  int order_;
  double lam_max_, lam_min_;
  bool verbose_ = false, debug_ = false, dynamic_ = false;
  const DDFT* ddft_ = nullptr;  // the DDFT whose matrix this is (dynamic_), for its use_squared_matrix_
};

// ------------------------------------------------------------- minimizer ---
// include/Minimizer.h:13-182, src/Minimizer.cpp.
class Minimizer {
 public:
  explicit Minimizer(DFT* dft) : dft_(dft) {}
  const DFT& getDFT() const { return *dft_; }
  virtual ~Minimizer() = default;
  void run(long maxSteps = -1);
  void resume(long maxSteps = -1);
  double getF() const { return F_; }
  int getCalls() const { return calls_; }
  void setForceTerminationCriterion(double v) { forceLimit_ = v; }
  double getForceTerminationCriterion() const { return forceLimit_; }
  // include/Minimizer.h:28-31: resume stops when N / V drops below minDensity; a fixed direction e reflects every
  // force, dF -= 2 (e . dF) e (src/Minimizer.cpp:77-84, 103-110)
  void setMinDensity(double m) { minDensity_ = m; }
  void set_fixed_direction(const DFT_Vec& fixed);
  void setVerbose(bool) {}
  virtual void reset() = 0;  // restart from the species density (src/Minimizer.cpp:30-34)
  virtual double get_convergence_monitor() const = 0;
  virtual double step() = 0;
  virtual void density(DFT_Vec& out) const = 0;  // current density (host copy)

 protected:
  void sync_species();
  DFT* dft_;
  double F_ = 0.0;
  int calls_ = 0;
  int step_counter_ = 0;
  double forceLimit_ = 0.1;
  double minDensity_ = -1;
  DFT_Vec fixed_direction_;
};

class fireMinimizer2 : public Minimizer {
 public:
  explicit fireMinimizer2(DFT* dft);
  ~fireMinimizer2() override;
  double step() override;
  // |vv| / V, src/Minimizer.h:105
  double get_convergence_monitor() const override { return rep_.monitor; }
  // max |dF/drho| / dV on the raw gradient (the criterion ||dF/drho||_inf of the configs)
  double get_raw_force_max() const { return rep_.fmax_raw; }
  double getRMS_Force() const { return rep_.rms_force; }
  void setTimeStep(double dt) { params_.dt = dt; }
  void setTimeStepMax(double dt) { params_.dt_max = dt; }
  void setAlphaStart(double a) { params_.alpha_start = a; }
  void setAlphaFac(double a) { params_.f_alf = a; }
  void setBacktrackFac(double a) { f_back_ = a; }  // stored only: the reference's step does not use it
  void reset() override;                           // fireMinimizer2::reset, src/Minimizer.cpp:149-158
  // src/Minimizer.cpp:259-268: every 10th step writes snapshot_s0.dat (the density) to the working directory;
  // n = 0 turns the snapshots off.  The step counter is per minimizer (a function-static in the reference).
  void set_snapshot_interval(int n) { snapshot_every_ = n; }
  // Runs FIRE steps on the GPU until max |dF/drho|/dV <= tol.
  int run_to_raw_tolerance(long maxSteps, double tol);
  // current density (host copy); with several species they follow one another, species 0 first
  void density(DFT_Vec& out) const override;

 private:
  void ensure();
  cdft_fire_params params_{};
  cdft_fire fire_ = nullptr;
  cdft_fire_report rep_{};
  int snapshot_every_ = 10;
  long snapshot_count_ = 0;
  double f_back_ = 0.8;
};

class DDFT : public Minimizer {
 public:
  // include/Minimizer.h:192; showGraphics has no effect here, central_differences selects the central-difference
  // g . x (src/DDFT.cpp:320-351) of the steps and of the dynamical matrix
  explicit DDFT(DFT* dft, bool showGraphics = true, bool central_differences = false);
  ~DDFT() override;
  // Dynamical_Matrix interface of the DDFT (include/Minimizer.h:211-220, src/DDFT.cpp:353-361 through
  // src/Dynamical_Matrix.cpp:34-70): result = g . (H v), applied twice and negated with use_squared_matrix_.
  void matrix_dot_v1(const DFT_Vec& v, DFT_Vec& result, void* param = nullptr) const;
  void set_use_squared_matrix_(bool b) { use_squared_matrix_ = b; }
  bool get_use_squared_matrix_() const { return use_squared_matrix_; }
  bool get_central_differences() const { return central_; }
  bool is_dynamic() const { return true; }
  double step() override;
  void reset() override {}  // Minimizer::reset only refreshes the FIRE aliases, which DDFT does not use
  void set_tolerence_fixed_point(double e) { tol_ = e; }
  void set_max_time_step(double t);
  void setTimeStep(double dt) { dt_ = dt; }
  double getTimeStep() const { return rep_.dt > 0 ? rep_.dt : dt_; }
  double get_time() const { return rep_.time; }
  double get_convergence_monitor() const override { return rep_.rhs_max; }
  double get_mass() const { return rep_.mass; }
  void density(DFT_Vec& out) const override;

 private:
  void ensure();
  double dt_ = 0.0, dtMax_ = 1.0, tol_ = 1e-4;
  cdft_ddft ddft_ = nullptr;
  bool use_squared_matrix_ = false, central_ = false;
  cdft_ddft_report rep_{};
};

// DDFT_IF_Open (include/Obsolete/Minimizer_old.h:288-325, mduran/src/DDFT_IF_Open.cpp:25-463): the density is
// held at `background` on the planes ix, iy or iz = 0 (an open system in contact with a reservoir) and the
// interior evolves through 3D DST-I transforms on the GPU.
class DDFT_IF_Open : public Minimizer {
 public:
  DDFT_IF_Open(DFT* dft, double background, bool showGraphics = false);
  ~DDFT_IF_Open() override;
  double step() override;
  void reset() override {}
  void set_tolerence_fixed_point(double e) { tol_ = e; }
  void set_max_time_step(double t) { dtMax_ = t; }
  void setTimeStep(double dt) { dt_ = dt; }
  double getTimeStep() const { return rep_.dt > 0 ? rep_.dt : dt_; }
  double get_time() const { return rep_.time; }
  double get_convergence_monitor() const override { return rep_.rhs_max; }
  void density(DFT_Vec& out) const override;

 private:
  void ensure();
  double background_, dt_ = 0.0, dtMax_ = 1.0, tol_ = 1e-4;
  cdft_ddft_open ddft_ = nullptr;
  cdft_ddft_open_report rep_{};
};

#endif  // CLASSICALDFT_AMD_H
