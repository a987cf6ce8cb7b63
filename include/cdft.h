/*
 * cdft.h - C ABI of the MI355X-native classicalDFT hot path.
 *
 * One "system" is one periodic grid with one species: the FMT hard-sphere
 * functional plus the optional ideal-gas, LJ mean-field (quadratic-force
 * interpolation) and external-field / chemical-potential terms.  A system
 * owns its device memory, its HIP stream and its resident weight spectra.
 *
 * All entry points return CDFT_OK (0) or a negative/positive status; the last
 * error message of the calling thread is available from cdft_last_error().
 * No C++ exception crosses this boundary.  Arrays are row-major with z
 * fastest: pos = z + Nz*(y + Ny*x), as in the reference Lattice::pos
 * (include/Lattice.h:120-126).
 *
 * Thread safety: one system per host thread at a time; distinct systems are
 * independent.
 */
#ifndef CDFT_H
#define CDFT_H

#ifdef __cplusplus
extern "C" {
#endif

enum cdft_status {
  CDFT_OK = 0,
  CDFT_ETA_TOO_LARGE = 1, /* eta > 1 somewhere: Eta_Too_Large_Exception, include/FMT.h:649-663 */
  CDFT_UPHILL_LIMIT = 2,  /* FIRE2: too many uphill steps, src/Minimizer.cpp:207 */
  CDFT_DDFT_STALLED = 3,  /* DDFT: the fixed point failed and dt / 10 is a change below the 1e-10 that
                             DDFT::change_timestep ignores (src/DDFT.cpp:184-186), so the reference's restart
                             loop would repeat forever; the step is abandoned with the state unchanged */
  CDFT_BAD_ARG = -1,
  CDFT_HIP_ERROR = -2,
  CDFT_NOT_SUPPORTED = -3
};

/* FMT models: include/FMT.h:139 (Rosenfeld), :262 (RSLT), :376 (esFMT A,B),
   :505 (WhiteBearI = esFMT(1,-1) + CS f3), :576 (WhiteBearII). */
enum cdft_model {
  CDFT_MODEL_NONE = -1,
  CDFT_ROSENFELD = 0,
  CDFT_RSLT = 1,
  CDFT_ESFMT = 2,
  CDFT_WHITEBEAR_I = 3,
  CDFT_WHITEBEAR_II = 4
};

enum cdft_precision {
  CDFT_FP64 = 0,  /* drop-in path: everything in double */
  CDFT_MIXED = 1  /* FFTs, spectra and convolutions in float; pointwise Phi,
                     ideal gas, reductions and solver state in double */
};

/* Pair potential of the mean field (include/Potential1.h) */
enum cdft_potential {
  CDFT_POT_LJ = 0,   /* Lennard-Jones, cut and shifted (:136-189) */
  CDFT_POT_TWF = 1,  /* ten Wolde-Frenkel (:195-257), range parameter alpha */
  CDFT_POT_WHDF = 2  /* Wang-Ramirez-Dobnikar-Frenkel (:262-324) */
};

/* Lattice kernel rule of the mean field (include/Interaction.h:130-431) */
enum cdft_mf_scheme {
  CDFT_MF_QF = 0,       /* Interaction_Interpolation_QF (default) */
  CDFT_MF_ZERO = 1,     /* Interaction_Interpolation_Zero */
  CDFT_MF_LE = 2,       /* Interaction_Interpolation_LE */
  CDFT_MF_QE = 3,       /* Interaction_Interpolation_QE */
  CDFT_MF_LF = 4,       /* Interaction_Interpolation_LF */
  CDFT_MF_GAUSS_E = 5,  /* Interaction_Gauss_E (mf_ngauss points) */
  CDFT_MF_GAUSS_F = 6   /* Interaction_Gauss_F (mf_ngauss points) */
};

enum cdft_flags {
  CDFT_DEVICE_PTRS = 1, /* rho / dF / vext are device pointers */
  CDFT_ACCUMULATE = 2,  /* add into dF instead of overwriting it */
  CDFT_FMT_ONLY = 4,    /* hard-sphere part only (cdft_system_eval_async / _time) */
  CDFT_NO_IDEAL = 8     /* no ideal-gas term: DFT::calculateFreeEnergyAndDerivatives(onlyFex = true) for
                           cdft_system_eval; DFT::full_hessian_ false for the Hessian-vector product */,
  CDFT_SQUARED_MATRIX = 16, /* log det / eigenvalue objective / dynamical matrix: apply the matrix twice
                               (use_squared_matrix_; negated for the dynamical matrix, is_dynamic) */
  CDFT_DYNAMIC_MATRIX = 32,  /* log det / eigenvalue objective: the matrix is the DDFT dynamical matrix
                               g . (H x) of cdft_system_dynamic_dot instead of the Hessian H */
  CDFT_CENTRAL_DIFFERENCES = 64  /* dynamical matrix: g . x with central differences (the DDFT was built with
                                    central_differences = true, src/DDFT.cpp:325, 345) */
};

/* Species boundary conditions (include/Species.h:24-25, applied in Species::endForceCalculation,
   src/Species.cpp:73-98).  Boundary points: ix, iy or iz = 0 (Lattice::is_boundary_point). */
enum cdft_boundary {
  CDFT_FIXED_BACKGROUND = 1,     /* dF = 0 on the boundary (setFixedBackground) */
  CDFT_HOMOGENEOUS_BOUNDARY = 2  /* boundary dF replaced by its mean (setHomogeneousBoundary) */
};

/* Generator of the FMT cell weights (FMT_Species::generateWeights, src/FMT_Species.cpp:310-496) */
enum cdft_weights_mode {
  CDFT_WEIGHTS_ACCURATE = 0,  /* default: the same integrals in a well-conditioned double-double form on the device,
                                 correct to the last bit of each cell weight */
  CDFT_WEIGHTS_REFERENCE = 1  /* the reference's own double expressions (asin forms, R += 1e-15 shift,
                                 src/FMT_Species.cpp:112-308) on the host, operation for operation: a reference run's
                                 weights wherever the libm and the floating-point contraction agree, with their
                                 rounding noise (up to ~1e-6 relative per cell at dx = 0.1, ~1e-4 at dx = 0.04) */
};

enum cdft_ao_force { CDFT_AO_FORCE_EXACT = 0, CDFT_AO_FORCE_REFERENCE = 1 };

typedef struct cdft_system_s* cdft_system;

typedef struct cdft_system_desc {
  int Nx, Ny, Nz;         /* grid; Nx, Ny, Nz/2 must be products of 2 and 3, <= 512; Nz even */
  double dx, dy, dz;      /* lattice spacings */
  int model;              /* enum cdft_model */
  double A, B;            /* esFMT coefficients (ignored for other models) */
  double hsd;             /* hard-sphere diameter */
  int precision;          /* enum cdft_precision */
  const double* fmt_weights; /* host, 11*N real-space cell weights in the
                                FMT_Species index order (include/Species.h:351-360),
                                or NULL to generate them on the device */
  int ideal;              /* include the ideal-gas term (src/DFT.cpp:388-394) */
  double mu;              /* beta*mu (src/Species.cpp:51-62) */
  const double* vext;     /* host, N values of beta*V_ext, or NULL */
  int mean_field;         /* include the LJ mean field (src/Interaction.cpp:76-115) */
  double lj_sigma, lj_eps, lj_rcut, kT; /* LJ with WCA split, include/Potential1.h:136-189 */
  const double* mf_kernel; /* host, N values of w_att (already / kT), or NULL to generate */
  /* mean-field potential and kernel rule; all zero = LJ with the WCA split, quadratic-force interpolation */
  int mf_potential;       /* enum cdft_potential (lj_sigma / lj_eps / lj_rcut are its sigma, eps, rcut) */
  double mf_alpha;        /* tWF range parameter (0: 50, the reference's default) */
  int mf_bh;              /* 1: Barker-Henderson split (Potential1::setBH) */
  double mf_r_att_min;    /* WCA continuation limit (Potential1::set_WCA_limit), 0 by default */
  int mf_scheme;          /* enum cdft_mf_scheme */
  int mf_ngauss;          /* Gauss-Legendre points per axis of CDFT_MF_GAUSS_E / _F (1..32) */
  /* Asakura-Oosawa polymer of radius ao_Rp at reservoir density ao_reservoir acting on the FMT species
     (FMT_AO_Species, include/Species.h:380-461; FP64 systems; ao_Rp = 0: none) */
  double ao_Rp, ao_reservoir;
  const double* ao_weights; /* host, 11*N polymer cell weights (diameter 2 Rp), or NULL to generate */
  /* Species constraints of the full evaluation (cdft_system_eval and the solvers; not cdft_fmt_eval):
     fixed_mass > 0 (Species::setFixedMass, src/Species.cpp:64-71, 99-110) rescales rho to that mass in place
     before each evaluation (the staged copy for host pointers), evaluates with mu = 0 and then projects
     dF -= mu dV with the multiplier mu = sum dF rho / fixed_mass (cdft_system_chem_potential); boundary:
     enum cdft_boundary flags.  Fixed mass and a fixed background exclude each other, as in the reference. */
  double fixed_mass;
  int boundary;
  /* factor on the mean-field kernel and a_vdw (Interaction_Base::scale_interaction, src/Interaction.cpp:54-59);
     0 means 1 */
  double mf_scale;
  /* how generated FMT (and AO polymer) cell weights are computed when fmt_weights / ao_weights is NULL:
     enum cdft_weights_mode */
  int weights_mode;
  /* AO polymer force, enum cdft_ao_force: CDFT_AO_FORCE_EXACT (0) is the derivative of the AO free energy;
     CDFT_AO_FORCE_REFERENCE counts each off-diagonal tensor measure once, as
     FMT_AO_Species::setFundamentalMeasures_AO does (include/Species.h:406-425: TI(j,k) and TI(k,j) name the
     same array, and the later write replaces the earlier one) */
  int ao_force;
} cdft_system_desc;

/* A pair potential for the host-side helpers below (fields as in cdft_system_desc). */
typedef struct cdft_potential_desc {
  int kind;               /* enum cdft_potential */
  double sigma, eps, rcut, alpha;
  int bh;
  double r_att_min;
} cdft_potential_desc;

const char* cdft_last_error(void);
int cdft_device_count(int* n);
/* The device on which the calling thread creates systems, mixtures and the other handles (each handle then stays
   on its device; the calling thread's device is the HIP current device).  No counterpart in the reference, which
   runs on the CPU; the device half of a context (device, stream) for callers without their own HIP runtime.
   CDFT_BAD_ARG for a device that does not exist. */
int cdft_set_device(int device);
int cdft_get_device(int* device);

/* Replaces the FMT_Species + FMT + Interaction + DFT construction
   (src/FMT_Species.cpp:95-108 weights; src/Interaction.cpp:32-52 kernel). */
int cdft_system_create(const cdft_system_desc* desc, cdft_system* out);
int cdft_system_destroy(cdft_system sys);

/* Replaces DFT::calculateFreeEnergyAndDerivatives(false) (src/DFT.cpp:360-425).
   parts (host, may be NULL) receives {F, F_id, F_hs, F_mf, F_ext}.  dF is
   dF/drho_i including dV (include/Species.h:55 convention). */
int cdft_system_eval(cdft_system sys, const double* rho, double* dF, double* parts, int flags);

/* Replaces FMT::calculateFreeEnergyAndDerivatives (src/FMT.cpp:363-427): only
   the hard-sphere part; returns F_hs (already x dV) in *F_hs and adds (or
   writes) the FMT dF. */
int cdft_fmt_eval(cdft_system sys, const double* rho, double* dF, double* F_hs, int flags);

/* Replaces DFT::matrix_dot_v_intern (src/DFT.cpp:435-488) for periodic systems:
   out = (d2 beta Omega / drho_i drho_j) v_j = dV v/rho (ideal gas; left out with
   CDFT_NO_IDEAL, i.e. full_hessian_ false) + the FMT part
   (FMT::add_second_derivative, src/FMT.cpp:437-567) + the mean field
   (Interaction_Base::add_second_derivative, src/Interaction.cpp:119-154).
   FP64 systems; batches take nb members of rho, v and out.  Flags:
   CDFT_DEVICE_PTRS, CDFT_ACCUMULATE, CDFT_NO_IDEAL.  CDFT_ETA_TOO_LARGE if eta > 1. */
int cdft_system_hessian_dot(cdft_system sys, const double* rho, const double* v, double* out, int flags);

/* No counterpart in the reference.  cdft_system_hessian_dot for callers that cannot allocate device memory at the
   call (code under HIP graph capture, tight serving loops): the same product on device pointers (CDFT_DEVICE_PTRS
   required; CDFT_NO_IDEAL, CDFT_ACCUMULATE as there), with its work spectra in a fixed 64 MiB array of the library
   on each GPU instead of system buffers allocated on first use.  Systems whose (nconv + 1) nb Nx Ny Zp complex
   doubles exceed 64 MiB (e.g. larger than 64^3 with White Bear + mean field) are refused with CDFT_BAD_ARG;
   fixed-background systems likewise.  Calls that overlap on one GPU (two host threads, or two streams) race on
   that array; each call synchronises the system's stream before it returns. */
int cdft_system_hessian_dot_noalloc(cdft_system sys, const double* rho, const double* v, double* out, int flags);

/* Replaces FMT::add_second_derivative (src/FMT.cpp:437-567): the hard-sphere
   part only; add into out with CDFT_ACCUMULATE as the reference does. */
int cdft_fmt_hessian_dot(cdft_system sys, const double* rho, const double* v, double* out, int flags);

/* Replaces DDFT::matrix_dot_v_intern (include/Minimizer.h:211-220, src/DDFT.cpp:353-361) as
   Dynamical_Matrix::matrix_dot_v1 applies it: out = g . (H v), H the Hessian at rho (cdft_system_hessian_dot) and
   g . x the DDFT mobility operator DDFT::g_dot_x (src/DDFT.cpp:320-351: forward differences, the density averaged
   over each link, periodic; zero on the boundary with a fixed background).  Flags: CDFT_DEVICE_PTRS, CDFT_NO_IDEAL,
   CDFT_SQUARED_MATRIX (out = -(gH)(gH) v, is_dynamic).  FP64 single systems. */
int cdft_system_dynamic_dot(cdft_system sys, const double* rho, const double* v, double* out, int flags);

/* Replaces Eigenvalues::calculate_gradients (include/Eigenvalues.h:44, src/Eigenvalues.cpp:46-70) with this
   system's DFT as the matrix: r = A x (Dynamical_Matrix::matrix_dot_v1: boundary entries of input and output zeroed
   with a fixed background; twice with CDFT_SQUARED_MATRIX), r += vshift (vshift . x) / 2 when vshift is not null,
   r *= (change_sign ? -1 : 1) scale; *f = x.r / x.x (the scaled Rayleigh quotient) and df = 2 (r - f x) / x.x.
   Flags: CDFT_DEVICE_PTRS (rho, x, vshift, df), CDFT_NO_IDEAL (full_hessian_ false), CDFT_SQUARED_MATRIX,
   CDFT_DYNAMIC_MATRIX (the DDFT's matrix: Eigenvalues(ddft)). */
int cdft_system_eigen_objective(cdft_system sys, const double* rho, const double* x, const double* vshift,
                                double scale, int change_sign, int flags, double* f, double* df);

/* Replaces Log_Det(matrix, order, lam_max, lam_min).calculate_log_det(seed, num_samples, has_zero_eigenvalue,
   variance) (include/Log_Det.h:22-28, src/Log_Det.cpp:22-166) with this system's DFT as the matrix: the
   stochastic Chebyshev estimate of log det of the Hessian at rho (Han, Malioutov, Avron and Shin 2017).  Probes are
   the reference's (std::mt19937(seed), uniform_int_distribution<>(0, 1) -> +-1, a random seed if seed <= 0); every
   Hessian-vector product runs on the device, as Log_Det::matrix_dot_v1 applies it (boundary entries of input and
   output zeroed with a fixed background).  Flags: CDFT_DEVICE_PTRS (rho), CDFT_NO_IDEAL (full_hessian_ false),
   CDFT_SQUARED_MATRIX, CDFT_DYNAMIC_MATRIX (Log_Det(ddft): each product negated, is_dynamic).  FP64 single
   systems; CDFT_ETA_TOO_LARGE if eta > 1. */
int cdft_system_log_det(cdft_system sys, const double* rho, int order, double lam_max, double lam_min, long seed,
                        int num_samples, int has_zero_eigenvalue, int flags, double* log_det, double* variance);

/* Asynchronous form on device pointers: enqueues one evaluation on the
   system's stream and returns without synchronising.  cdft_system_fetch
   waits for the stream and returns the parts of the last evaluation
   (CDFT_ETA_TOO_LARGE if any evaluation since the previous fetch saw eta > 1). */
int cdft_system_eval_async(cdft_system sys, const double* rho_d, double* dF_d, int flags);
int cdft_system_fetch(cdft_system sys, double* parts);
/* Enqueues `iters` asynchronous evaluations between two hipEvents recorded on
   the system's stream; *ms = elapsed time of the whole batch. */
int cdft_system_time(cdft_system sys, const double* rho_d, double* dF_d, int flags, int iters, double* ms);

/* Bulk thermodynamics of the system's functional at uniform density rho:
   beta f = rho ln rho - rho + Phi(uniform) + a_vdw rho^2 / 2 (per volume) and
   beta mu = ln rho + mu_ex + a_vdw rho.  Replaces DFT::mu_times_beta and
   fhelmholtz_times_beta_over_volume (src/DFT.cpp:58-113; FMT::BulkMuex /
   BulkFex src/FMT.cpp:616-709; Interaction_Base::Mu src/Interaction.cpp:61-72). */
int cdft_system_bulk(cdft_system sys, double rho, double* f, double* mu);
/* Species::getChemPotential after the last fetched evaluation: the fixed-mass multiplier
   (src/Species.cpp:99-110) of each member, or the desc mu(s) without a fixed mass. */
int cdft_system_chem_potential(cdft_system sys, double* mu);
/* Points-file mean-field kernel: Interaction::generateWeights (src/Interaction.cpp:672-860).  points: npoints
   rows of (ux, uy, uz, weight) with unit directions and the weights already divided by npoints, as the reference
   reads its points file (:744-764); nradial Gauss-Legendre nodes on [0, rcut] (0: the reference's 512); the 6
   axis permutations and 8 reflections of every direction.  w_out (host, N values) receives the kernel BEFORE the
   division by kT (what the reference's weight cache stores); pass w_out / kT as desc mf_kernel.  Atomic
   accumulation on the GPU: cells agree with a serial sum to rounding, not bit for bit. */
int cdft_mf_kernel_points(int Nx, int Ny, int Nz, double dx, double dy, double dz, const cdft_potential_desc* pot,
                          const double* points, int npoints, int nradial, double* w_out);
/* One fixed mass per batch member (nb values > 0) for a system created with fixed_mass > 0 (the whole-DFT
   batch with per-system mu, V_ext and fixed mass). */
int cdft_system_set_fixed_masses(cdft_system sys, const double* masses);

/* Hard-sphere diameter of the WCA-split Lennard-Jones potential (cut and
   shifted at rcut), d = int_0^rmin (1 - exp(-V0(r)/kT)) dr with
   V0 = V - V(rmin) for r < rmin, rmin = 2^(1/6) sigma.  Replaces
   Potential1::getHSD (include/Potential1.h:73-83; GSL qags there, adaptive
   Gauss-Kronrod 7/15 to 1e-13 absolute here).  Host-only. */
int cdft_lj_hsd(double sigma, double eps, double rcut, double kT, double* hsd);

/* Potential1::getHSD (include/Potential1.h:73-83) for any potential and split:
   hard core + int_hc^rlimit (1 - exp(-V0(r)/kT)) dr, rlimit = rmin (WCA) or
   r0 (BH).  Adaptive Gauss-Kronrod 7/15 to 1e-13.  Host-only. */
int cdft_potential_hsd(const cdft_potential_desc* p, double kT, double* hsd);
/* Potential1::getVDW_Parameter (include/Potential1.h:85-91): (2 pi / kT) int r^2
   Watt(r) dr over [r_att_min, rmin] + [rmin, rcut] (WCA) or [r0, rcut] (BH). */
int cdft_potential_vdw(const cdft_potential_desc* p, double kT, double* a);
/* Potential1::Watt at n distances (host). */
int cdft_potential_watt(const cdft_potential_desc* p, int n, const double* r, double* w);

/* ---------------------------------------------------------------- FIRE2 ---
   fireMinimizer2 (include/Minimizer.h:86-182, src/Minimizer.cpp:117-327) on
   the alias variables of FMT_Species (src/FMT_Species.cpp:29-93).  The
   iteration state lives on the device; the control logic is the reference's. */
typedef struct cdft_fire_s* cdft_fire;
typedef struct cdft_fire_params {
  double dt, dt_max, dt_min, alpha_start, f_dec, f_inc, f_alf;
  int N_delay, N_P_negative_max, initial_delay;
} cdft_fire_params;
typedef struct cdft_fire_report {
  double F;          /* free energy after the step */
  double fmax_alias; /* max |dF/dx| / dV on the alias variables (the reference's fmax_) */
  double fmax_raw;   /* max |dF/drho| / dV (Species.h:55 convention) */
  double dt, alpha, vv, rms_force;
  double monitor;    /* |vv| / V: fireMinimizer2::get_convergence_monitor */
  int steps, calls;
} cdft_fire_report;
int cdft_fire_default_params(cdft_fire_params* p);
/* Constructor: alias from rho0 (host unless flags has CDFT_DEVICE_PTRS), first
   getDF_DX, v = 0.  params may be NULL (reference defaults). */
int cdft_fire_create(cdft_system sys, const double* rho0, int flags, const cdft_fire_params* params, cdft_fire* out);
/* Minimizer::set_fixed_direction (include/Minimizer.h:31, used in getDF_DX, src/Minimizer.cpp:103-110): every
   force is reflected, dF -= 2 (e . dF) e, before the alias conversion; dir = NULL clears it. */
int cdft_fire_set_fixed_direction(cdft_fire f, const double* dir, int flags);
/* fireMinimizer2::reset (src/Minimizer.cpp:30-34, 149-158): restart from rho (aliases, one evaluation, v = 0),
   keeping dt, alpha and the step counters. */
int cdft_fire_reset(cdft_fire f, const double* rho, int flags);
/* One fireMinimizer2::step.  Returns CDFT_OK, CDFT_ETA_TOO_LARGE (the step
   back-tracked: x, v/2 and dF restored, dt halved) or CDFT_UPHILL_LIMIT. */
int cdft_fire_step(cdft_fire f, cdft_fire_report* rep);
/* Steps until fmax_raw <= tol_raw or max_steps (back-tracks included). */
int cdft_fire_run(cdft_fire f, int max_steps, double tol_raw, cdft_fire_report* rep);
int cdft_fire_get(cdft_fire f, double* rho, double* dF_raw, double* x, double* v);
int cdft_fire_destroy(cdft_fire f);

/* ----------------------------------------------------------------- DDFT ---
   Integrating-factor DDFT (src/DDFT.cpp:42-351), single species. */
typedef struct cdft_ddft_s* cdft_ddft;
typedef struct cdft_ddft_report {
  double F, dt, time, rhs_max, mass;
  int calls, restarts;
} cdft_ddft_report;
/* dt <= 0 selects the reference default 1e-4 dx^2 (src/DDFT.cpp:46); dt_max
   <= 0 selects 1; tol = 0 selects 1e-4 (include/Minimizer.h:239), any other value is used as given
   (DDFT::set_tolerence_fixed_point).  cdft_ddft_step returns CDFT_DDFT_STALLED where the reference's restart
   loop would not end. */
int cdft_ddft_create(cdft_system sys, const double* rho0, double dt, double dt_max, double tol, cdft_ddft* out);
int cdft_ddft_step(cdft_ddft d, cdft_ddft_report* rep);
int cdft_ddft_set_max_time_step(cdft_ddft d, double dt_max);
/* The third argument of DDFT(DFT*, bool showGraphics, bool central_differences) (include/Minimizer.h:192): with
   on != 0 the excess RHS uses DDFT::g_dot_x with central differences (src/DDFT.cpp:320-351: dF's neighbours two
   sites away, the link density without the site's own, D = 1/(4 dx^2)).  Off by default, as in the reference. */
int cdft_ddft_set_central_differences(cdft_ddft d, int on);
int cdft_ddft_get_density(cdft_ddft d, double* rho);
int cdft_ddft_destroy(cdft_ddft d);

/* Open-boundary DDFT: DDFT_IF_Open (mduran/src/DDFT_IF_Open.cpp:25-463, declared in
   include/Obsolete/Minimizer_old.h:288-325).  The density is held at `background` on the planes ix, iy or
   iz = 0; the interior evolves with the integrating factor of the Dirichlet Laplacian through 3D DST-I
   transforms.  step = DDFT_IF_Open::step (step_string: at most 20 fixed-point iterations, dt halved on a
   restart and restored for the next step).  Nx, Ny, Nz: 2N a product of 2 and 3, <= 1024.  Systems
   without species constraints; dt = 0 gives 1e-4 dx^2, dt_max <= 0 and tol = 0 the defaults 1 and 1e-4. */
typedef struct cdft_ddft_open_s* cdft_ddft_open;
typedef struct cdft_ddft_open_report {
  double F, dt, time;
  double rhs_max;  /* max |g . dF| / dV^2 of the step's starting density (DDFT::get_convergence_monitor) */
  int calls, restarts, evaluations;
} cdft_ddft_open_report;
int cdft_ddft_open_create(cdft_system sys, const double* rho0, double background, double dt, double dt_max, double tol,
                          cdft_ddft_open* out);
int cdft_ddft_open_step(cdft_ddft_open d, cdft_ddft_open_report* rep);
int cdft_ddft_open_get_density(cdft_ddft_open d, double* rho);
int cdft_ddft_open_destroy(cdft_ddft_open d);
/* 3D DST-I (FFTW RODFT00) of a compact (n0, n1, n2) host array, in place, on the GPU. */
int cdft_dst1_3d(int n0, int n1, int n2, double* a);

/* Device memory plumbing for callers without their own HIP runtime. */
int cdft_malloc(void** p, unsigned long long bytes);
int cdft_free(void* p);
int cdft_memcpy_h2d(void* dst, const void* src, unsigned long long bytes);
int cdft_memcpy_d2h(void* dst, const void* src, unsigned long long bytes);

/* Introspection / plumbing. */
int cdft_system_stream(cdft_system sys, void** hip_stream);
/* The stream half of that context: the system's later work (evaluations, solvers, graphs they launch) runs on
   hip_stream (a hipStream_t of the system's device, kept alive by the caller while the system uses it), so that it
   orders with the caller's own work; NULL returns to the system's own stream.  The work already enqueued finishes
   first.  No counterpart in the reference.  Not for the member systems of a mixture. */
int cdft_system_set_stream(cdft_system sys, void* hip_stream);
int cdft_system_a_vdw(cdft_system sys, double* a_vdw); /* Interaction_Base::getVDWParameter */
/* The weighted-density arrays the pipeline stores (11 for tensor models, 10 when the weights satisfy the trace
   identity and T_zz is rebuilt per site, 5 without tensor) and those it convolves (plus the mean-field kernel).
   No counterpart in the reference (its FMT_Species always stores 11, include/Species.h:351-360); for callers that
   size buffers or count bytes. */
int cdft_system_layout(cdft_system sys, int* stored_arrays, int* convolved_arrays);
/* ---- Hard-sphere mixtures ----
   Several FMT species of different diameters on one lattice: FMT::calculateFreeEnergyAndDerivatives with more
   than one FMT_Species in allSpecies (src/FMT.cpp:363-427).  The fundamental measures sum every species' weighted
   densities scaled by its diameter (FMT::getWeightedDensities, src/FMT.cpp:78-111), Phi is taken once per site,
   and each species collapses dPhi with its own diameter (FMT_Species::set_fundamental_measure_derivatives,
   include/Species.h:310-329) before its back-convolution.  FP64; Nx, Ny and Nz/2 must be register-FFT lengths
   (4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 256). */
typedef struct cdft_mixture_s* cdft_mixture;

typedef struct cdft_mixture_desc {
  int Nx, Ny, Nz;
  double dx, dy, dz;
  int model;                         /* enum cdft_model (an FMT model) */
  double A, B;                       /* esFMT coefficients */
  int nspecies;                      /* >= 1 */
  const double* hsd;                 /* host, nspecies hard-sphere diameters (FMT_Species::getHSD) */
  const double* const* fmt_weights;  /* host, NULL, or nspecies pointers to 11*N cell weights (NULL entries are
                                        generated on the device, as for a single species) */
  int weights_mode;                  /* enum cdft_weights_mode */
  /* the rest of DFT::calculateFreeEnergyAndDerivatives per species (src/DFT.cpp:377-425, src/Species.cpp:51-62),
     used by cdft_mixture_eval and the mixture's FIRE: */
  int ideal;                         /* include each species' ideal-gas term */
  const double* mu;                  /* host, NULL or nspecies values of beta*mu */
  const double* const* vext;         /* host, NULL or nspecies pointers (NULL entries: no field) to N values of
                                        beta*V_ext */
  /* Species constraints per species (include/Species.h:23-35, src/Species.cpp:64-112), as the desc fields of one
     system: fixed_mass[s] > 0 rescales rho_s to that mass before each evaluation and projects its force with the
     multiplier that cdft_mixture_chem_potential returns; boundary[s]: enum cdft_boundary flags. */
  const double* fixed_mass;          /* host, NULL or nspecies values (0: free) */
  const int* boundary;               /* host, NULL or nspecies flag sets */
} cdft_mixture_desc;

/* Replaces the construction of the species' FMT_Species weights (src/FMT_Species.cpp:95-108, 310-496) for a
   mixture. */
int cdft_mixture_create(const cdft_mixture_desc* desc, cdft_mixture* out);
int cdft_mixture_destroy(cdft_mixture mix);
int cdft_mixture_species(cdft_mixture mix, int* nspecies);
/* Replaces FMT::calculateFreeEnergyAndDerivatives(vector<Species*>&) over several FMT_Species
   (src/FMT.cpp:363-427): rho[s], dF[s] are nspecies arrays of N values (host, or device with CDFT_DEVICE_PTRS);
   *F_hs receives the hard-sphere free energy (x dV); dF[s] receives dF_hs/drho_s (x dV), written or, with
   CDFT_ACCUMULATE, added to (Species::addToForce).  CDFT_ETA_TOO_LARGE as cdft_fmt_eval. */
int cdft_mixture_fmt_eval(cdft_mixture mix, const double* const* rho, double* const* dF, double* F_hs, int flags);
/* Replaces DFT::calculateFreeEnergyAndDerivatives(onlyFex) with several species (src/DFT.cpp:360-425): per species
   the ideal gas (unless CDFT_NO_IDEAL), dV sum V_ext rho - mu N, and the mixture's hard-sphere term.  parts (host,
   may be NULL) receives {F, F_id, F_hs, F_mf = 0, F_ext} summed over species.  The reference's mean field between
   species is not implemented there (src/Interaction.cpp:110) and is not here. */
int cdft_mixture_eval(cdft_mixture mix, const double* const* rho, double* const* dF, double* parts, int flags);
/* Species::getChemPotential of every species after the last cdft_mixture_eval (or FIRE step): the fixed-mass
   multiplier of a species with a fixed mass, its mu otherwise (nspecies values). */
int cdft_mixture_chem_potential(cdft_mixture mix, double* mu);
/* fireMinimizer2 over all species of the mixture (src/Minimizer.cpp:160-327: one FIRE2 state, P, norms and the
   mixing over the species' alias variables together), on cdft_mixture_eval.  The returned handle is driven by
   cdft_fire_step / _run / _destroy; cdft_fire_get and cdft_fire_reset take nspecies x N values (species after
   species).  rho0: nspecies pointers to N values (host, or device with CDFT_DEVICE_PTRS). */
int cdft_mixture_fire_create(cdft_mixture mix, const double* const* rho0, int flags, const cdft_fire_params* params,
                             cdft_fire* out);

/* Copies the device-generated real-space FMT weights (11*N) or MF kernel (N). */
int cdft_system_get_weights(cdft_system sys, double* w11);
int cdft_system_get_mf_kernel(cdft_system sys, double* w);

/* Standalone 3D FFTs with the FFTW conventions of DFT_FFT (include/DFT_LinAlg.h:210-211):
   unnormalised r2c (N real -> Nx*Ny*(Nz/2+1) complex interleaved) and c2r.
   Host pointers; used to check the transform against numpy / the oracle. */
int cdft_fft_r2c(int Nx, int Ny, int Nz, int precision, const double* in, double* out);
int cdft_fft_c2r(int Nx, int Ny, int Nz, int precision, const double* in, double* out);

/* Per-stage timing of the last cdft_system_eval with timing enabled: up to
   16 entries of {kernel id, milliseconds}, one per stage name (a stage that
   runs once per x-slab is summed over the slabs).  enable = 1 records
   hipEvents between stages (adds no synchronisation inside the pipeline). */
int cdft_system_set_timing(cdft_system sys, int enable);
int cdft_system_get_timing(cdft_system sys, int* n, double* ms, const char** names);

/* ------------------------------------------------------------- batches ---
   No counterpart in the reference, which holds one system per DFT object;
   for callers that step many independent systems at once (parameter scans,
   training-data generation).
   A batch is one cdft_system with nb members of the same grid and functional
   (desc: shape, spacing, model, hsd, precision, ideal / mean-field / LJ
   fields; desc->mu and desc->vext are ignored) with their own mu[nb] and
   V_ext (nb x N, member-major, or NULL).  The weight spectra are shared and
   every kernel of the evaluation runs all members in one launch.
   cdft_system_eval / _eval_async / _fetch / _time and cdft_fmt_eval on a
   batch take nb x N densities and forces and return nb x 5 parts (nb F_hs).
   Needs Nx, Ny and Nz/2 in {4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 256}. */
int cdft_system_batch_create(const cdft_system_desc* desc, int nb, const double* mu, const double* vext,
                             cdft_system* out);

/* Tile autotuner: times `reps` evaluations of a uniform density for each
   kernel variant (register or LDS Stockham transforms, columns per block of
   the x / y passes and of both convolutions, rows per block of the pointwise
   kernel), keeps the fastest on `sys` and caches it per (shape, precision,
   functional, member count) for the process.  Overwrites the system's own
   density / force buffers (not the caller's). */
int cdft_system_autotune(cdft_system sys, int reps, double* best_ms);

/* FIRE2 on every member of a batch with the control logic on the device
   (same step as cdft_fire_step, src/Minimizer.cpp:160-327, per member).
   create: evaluates the start densities rho0 (nb x N, host).  run: enqueues
   `steps` steps of every member that has not converged (max_i |dF/drho_i| /
   dV <= tol_raw) or failed; use_graph = 1 captures the `steps` steps once
   into a hipGraph and launches it (recaptured when steps or tol change).
   run does not synchronise; status does, and reports per member (status 0
   running, 1 converged, 2 uphill limit) and the number still running. */
typedef struct cdft_fire_batch_s* cdft_fire_batch;
int cdft_fire_batch_create(const cdft_system_desc* desc, int nb, const double* mu, const double* vext,
                           const double* rho0, const cdft_fire_params* params, cdft_fire_batch* out);
int cdft_fire_batch_run(cdft_fire_batch b, int steps, double tol_raw, int use_graph);
/* Switches the evaluations to FP64 or MIXED (the other system is created on
   first use): a precision ladder, mixed-precision steps until the FP32
   gradient noise (~1e-6 in max |dF/drho|) then FP64 to the final tolerance.
   Re-evaluates the current densities and restarts the FIRE2 control of every
   running member (v = 0, dt, dt_max, alpha and counters from the parameters):
   a noisy mixed-precision end game can shrink dt_max by back-tracking.
   Members marked converged run again. */
int cdft_fire_batch_set_precision(cdft_fire_batch b, int precision);
int cdft_fire_batch_status(cdft_fire_batch b, cdft_fire_report* reports, int* status, int* running);
int cdft_fire_batch_get(cdft_fire_batch b, int member, double* rho);
/* The FIRE2 state of one member, as cdft_fire_get (any pointer may be null): density, raw gradient dF/drho of the
   last evaluation, alias coordinates and velocities (fireMinimizer2's x_ and v_, include/Minimizer.h:105-160). */
int cdft_fire_batch_state(cdft_fire_batch b, int member, double* rho, double* dF_raw, double* x, double* v);
int cdft_fire_batch_system(cdft_fire_batch b, cdft_system* sys);
void cdft_fire_batch_destroy(cdft_fire_batch b);
/* No reference counterpart (the reference runs one fireMinimizer2 per system, src/Minimizer.cpp:160-271): several
   batches, e.g. one per grid size, stepped together.  cdft_fire_group_run advances every batch by `steps` FIRE2 steps
   with one hipGraph launch whose branches (one per batch, each on the batch's stream) run concurrently; members stop
   at tol_raw as in cdft_fire_batch_run.  The batches stay owned by the caller and usable on their own
   (status, get, set_precision: the group re-captures after a precision change); destroy the group first. */
typedef struct cdft_fire_group_s* cdft_fire_group;
int cdft_fire_group_create(cdft_fire_batch* batches, int n, cdft_fire_group* out);
int cdft_fire_group_run(cdft_fire_group g, int steps, double tol_raw);
void cdft_fire_group_destroy(cdft_fire_group g);

/* ------------------------------------------------------- learned warm start ---
   No counterpart in the reference.  A 3-level 3D U-Net on periodic D^3 grids
   (D a multiple of 32): 3x3x3 convolutions with 16 / 32 / 64 channels, ReLU,
   2x average pooling, nearest upsampling with skip concatenation, a 1x1 head;
   input channels beta V_ext / 3 and (mu + 3) / 2, output log rho.  BF16
   weights and activations, FP32 accumulation (MFMA), FP32 master weights and
   Adam.  Data are device arrays: V (n_samples x D^3 floats), mu (n_samples),
   targets log rho (n_samples x D^3); idx (host, nb ints) picks the batch.
   out (device, nb x D^3 floats).  grad returns the loss (mean squared error of
   log rho) and the FP32 gradient of all parameters (host, num_params floats)
   without updating them; train_step also applies one Adam step. */
typedef struct cdft_unet_s* cdft_unet;
int cdft_unet_create(int D, unsigned long long seed, cdft_unet* out);
void cdft_unet_destroy(cdft_unet u);
int cdft_unet_num_params(cdft_unet u, long long* n);
int cdft_unet_get_params(cdft_unet u, float* host);
int cdft_unet_set_params(cdft_unet u, const float* host);
int cdft_unet_forward(cdft_unet u, const float* V, const float* mu, const int* idx, int nb, float* out);
int cdft_unet_grad(cdft_unet u, const float* V, const float* mu, const float* target, const int* idx, int nb,
                   float* loss, float* grad);
int cdft_unet_train_step(cdft_unet u, const float* V, const float* mu, const float* target, const int* idx, int nb,
                         float lr, float* loss);

#ifdef __cplusplus
}
#endif
#endif /* CDFT_H */
