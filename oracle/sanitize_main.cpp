// Host sanitizer driver for the CPU oracle (test infrastructure, like oracle.cpp itself).
//
// Built together with oracle.cpp under AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitizers.py (g++ -fsanitize=address,undefined); it walks every major entry point of the oracle on
// small grids (3D FFTs, weight generation, the five FMT models, a two-species mixture with its DFT and FIRE, the whole DFT with mean field, external field and
// constraints, the Hessian, AO, FIRE2, DDFT, open DDFT, log-det and eigenvalue objectives) so that an out-of-bounds
// access, a use after free or undefined arithmetic anywhere on those paths stops the run.  Prints one line per
// stage and "sanitize: ok" at the end.
#include <cmath>
#include <cstdio>
#include <vector>

extern "C" {
struct orc_fmt;
struct orc_dft;
struct orc_fire;
struct orc_ddft;
struct orc_open;
void orc_set_threads(int pointwise, int fft);
void orc_r2c3d(int Nx, int Ny, int Nz, const double* in, double* out);
void orc_c2r3d(int Nx, int Ny, int Nz, const double* in, double* out);
int orc_fmt_weights(int Nx, int Ny, int Nz, double dx, double dy, double dz, double hsd, double* w, int mode);
int orc_fmt_eval(int kind, double A, double B, int Nx, int Ny, int Nz, double dx, double dy, double dz, double hsd,
                 const double* weights, const double* rho, double* dF, double* F, double* phi_out);
int orc_fmt_eval_mix(int kind, double A, double B, int Nx, int Ny, int Nz, double dx, double dy, double dz, int S,
                     const double* hsd, const double* weights, const double* rho, double* dF, double* F);
orc_dft* orc_dft_create(int kind, double A, double B, int Nx, int Ny, int Nz, double dx, double dy, double dz,
                        double hsd, int has_fmt, int ideal, int has_mf, double sigma, double eps, double rcut, double kT,
                        double mu, const double* vext);
void orc_dft_destroy(orc_dft* h);
int orc_dft_eval(orc_dft* h, double* rho, double* dF, double* parts);
int orc_dft_hess(orc_dft* h, const double* rho, const double* v, double* out, int full);
void orc_dft_set_ao(orc_dft* h, double Rp, double rho_p, const double* ao_weights, int exact);
int orc_dft_set_constraints(orc_dft* h, double fixed_mass, int bnd);
int orc_dft_log_det(orc_dft* h, const double* rho, int order, double lam_max, double lam_min, long seed,
                    int num_samples, int has_zero, int full, int squared, double* log_det, double* variance);
double orc_dft_eig_objective(orc_dft* h, const double* rho, const double* x, const double* vshift, double scale,
                             int change_sign, int full, int squared, double* df);
struct orc_mix;
orc_mix* orc_mix_create(int kind, double A, double B, int Nx, int Ny, int Nz, double dx, double dy, double dz, int S,
                        const double* hsd, const double* weights, int ideal, const double* mu, const double* vext,
                        const double* fixed_mass, const int* bnd);
void orc_mix_destroy(orc_mix* h);
int orc_mix_eval(orc_mix* h, const double* rho, double* dF, double* parts);
orc_fire* orc_mix_fire_create(orc_mix* h, const double* rho0, double dt, double dt_max);
orc_fire* orc_fire_create(orc_dft* h, const double* rho0, double dt, double dt_max);
int orc_fire_step(orc_fire* h, double* info);
void orc_fire_destroy(orc_fire* h);
orc_ddft* orc_ddft_create(orc_dft* h, const double* rho0, double dt, double dtMax, double tol);
int orc_ddft_step(orc_ddft* h, double* info);
void orc_ddft_destroy(orc_ddft* h);
orc_open* orc_open_create(orc_dft* h, const double* rho0, double background, double dt, double dtMax, double tol);
int orc_open_step(orc_open* h, double* info);
void orc_open_destroy(orc_open* h);
}

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("sanitize: FAILED %s\n", what);
    ++failures;
  }
}

std::vector<double> density(int n, double rho0, unsigned seed) {
  std::vector<double> r(n);
  unsigned x = seed;
  for (int i = 0; i < n; ++i) {
    x = x * 1664525u + 1013904223u;
    r[i] = rho0 * (1.0 + 0.1 * ((x >> 8) / 16777216.0 - 0.5));
  }
  return r;
}

}  // namespace

int main() {
  orc_set_threads(2, 2);
  const int Nx = 12, Ny = 10, Nz = 14, N = Nx * Ny * Nz, Nk = Nx * Ny * (Nz / 2 + 1);
  const double dx = 0.1;
  std::vector<double> rho = density(N, 0.5, 7), dF(N), spec(2 * Nk), back(N);

  orc_r2c3d(Nx, Ny, Nz, rho.data(), spec.data());
  orc_c2r3d(Nx, Ny, Nz, spec.data(), back.data());
  double err = 0.0;
  for (int i = 0; i < N; ++i) err = std::fmax(err, std::fabs(back[i] / N - rho[i]));
  expect(err < 1e-12, "fft round trip");
  std::printf("fft: round trip %.2e\n", err);

  std::vector<double> w(11L * N);
  for (int mode = 0; mode < 3; ++mode) expect(orc_fmt_weights(Nx, Ny, Nz, dx, dx, dx, 0.6, w.data(), mode) == 0, "weights");
  std::printf("weights: three modes\n");

  for (int kind = 0; kind < 5; ++kind) {
    double F = 0.0;
    std::vector<double> phi(N);
    expect(orc_fmt_eval(kind, 1.3, -0.4, Nx, Ny, Nz, dx, dx, dx, 0.6, w.data(), rho.data(), dF.data(), &F,
                        phi.data()) == 0,
           "fmt eval");
    expect(std::isfinite(F), "fmt F finite");
  }
  std::printf("fmt: five models\n");

  {  // a two-species mixture (weights generated per species)
    const double hsd2[2] = {0.6, 0.45};
    std::vector<double> r2(2 * (size_t)N), d2(2 * (size_t)N);
    for (int i = 0; i < N; ++i) {
      r2[i] = 0.6 * rho[i];
      r2[N + i] = 0.4 * rho[i];
    }
    double F2 = 0.0;
    expect(orc_fmt_eval_mix(3, 1.0, -1.0, Nx, Ny, Nz, dx, dx, dx, 2, hsd2, nullptr, r2.data(), d2.data(), &F2) == 0,
           "mixture eval");
    expect(std::isfinite(F2), "mixture F finite");
    std::printf("fmt: two-species mixture\n");
    // the mixture's DFT (ideal gas, mu, a field on the second species) and five FIRE2 steps over both species
    const double mu2[2] = {0.3, -0.2};
    std::vector<double> v2(2 * (size_t)N, 0.0), parts(5);
    for (int i = 0; i < N; ++i) v2[N + i] = 0.2 * std::cos(0.3 * i);
    const double fm2[2] = {0.0, 0.4 * Nx * Ny * Nz * dx * dx * dx * 0.5};  // the second species' mass fixed
    orc_mix* mx = orc_mix_create(3, 1.0, -1.0, Nx, Ny, Nz, dx, dx, dx, 2, hsd2, nullptr, 1, mu2, v2.data(), fm2, nullptr);
    expect(orc_mix_eval(mx, r2.data(), d2.data(), parts.data()) == 0, "mixture DFT eval");
    expect(std::isfinite(parts[0]), "mixture DFT F finite");
    orc_fire* fm = orc_mix_fire_create(mx, r2.data(), 1e-3, 1e-2);
    double info[8];
    for (int k = 0; k < 5; ++k) expect(orc_fire_step(fm, info) != 2, "mixture fire step");
    expect(std::isfinite(info[0]), "mixture fire F finite");
    orc_fire_destroy(fm);
    orc_mix_destroy(mx);
    std::printf("fmt: two-species mixture DFT and FIRE\n");
  }

  std::vector<double> vext(N);
  for (int i = 0; i < N; ++i) vext[i] = 0.3 * std::sin(0.1 * i);
  orc_dft* d = orc_dft_create(3, 1.0, -1.0, Nx, Ny, Nz, dx, dx, dx, 0.6, 2, 1, 1, 0.3, 1.0, 0.45, 0.9, -0.5,
                              vext.data());
  double parts[5];
  expect(orc_dft_eval(d, rho.data(), dF.data(), parts) == 0, "dft eval");
  std::vector<double> v = density(N, 1.0, 11), hv(N);
  expect(orc_dft_hess(d, rho.data(), v.data(), hv.data(), 1) == 0, "hessian");
  double ld = 0.0, var = 0.0;
  orc_dft_log_det(d, rho.data(), 6, 50.0, 0.1, 3, 2, 0, 1, 0, &ld, &var);
  std::vector<double> g(N);
  const double obj = orc_dft_eig_objective(d, rho.data(), v.data(), nullptr, 1.0, 0, 1, 0, g.data());
  expect(std::isfinite(obj), "eigen objective");
  std::printf("dft: eval, hessian, log det, eigen objective\n");

  orc_fire* f = orc_fire_create(d, rho.data(), 1e-3, 1e-2);
  double info[8];
  for (int k = 0; k < 5; ++k) orc_fire_step(f, info);
  orc_fire_destroy(f);
  orc_ddft* dd = orc_ddft_create(d, rho.data(), 1e-4, 1e-3, 1e-4);
  for (int k = 0; k < 2; ++k) orc_ddft_step(dd, info);
  orc_ddft_destroy(dd);
  orc_open* op = orc_open_create(d, rho.data(), 0.5, 1e-4, 1e-3, 1e-4);
  for (int k = 0; k < 2; ++k) orc_open_step(op, info);
  orc_open_destroy(op);
  std::printf("solvers: fire, ddft, open ddft\n");

  expect(orc_dft_set_constraints(d, 0.5 * N * dx * dx * dx, 0) == 0, "fixed mass");
  expect(orc_dft_eval(d, rho.data(), dF.data(), parts) == 0, "dft eval with fixed mass");
  expect(orc_dft_set_constraints(d, 0.0, 3) == 0, "boundary constraints");
  expect(orc_dft_eval(d, rho.data(), dF.data(), parts) == 0, "dft eval with boundary constraints");
  orc_dft_destroy(d);

  orc_dft* a = orc_dft_create(3, 1.0, -1.0, Nx, Ny, Nz, dx, dx, dx, 0.6, 2, 1, 0, 1.0, 1.0, 2.5, 1.0, -0.5, nullptr);
  orc_dft_set_ao(a, 0.2, 0.4, nullptr, 0);
  expect(orc_dft_eval(a, rho.data(), dF.data(), parts) == 0, "ao eval");
  orc_dft_destroy(a);
  std::printf("constraints, ao\n");

  if (failures) return 1;
  std::printf("sanitize: ok\n");
  return 0;
}
