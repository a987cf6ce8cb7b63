// Pair potentials and the lattice mean-field kernel rules (host and device).
//
// Potentials: include/Potential1.h (Potential1 :23-133, LJ :136-189, tWF
// :195-257, WHDF :262-324): the cut-and-shifted potential V, its attractive
// tail Watt (WCA split with continuation inside r_min, or the BH split at the
// zero of V), the repulsive part V0.
//
// Kernel rules: Interaction_Interpolation_{Zero,LE,QE,LF,QF}
// (include/Interaction.h:293-431, generateWeight src/Interaction.cpp:522-534)
// are tensor products of a 1D rule (points p_i, weights v_i) applied to
// Watt2(|S + p|^2); Interaction_Gauss_{E,F} (src/Interaction.cpp:443-520)
// integrate cubic-spline (E, /216) or trilinear (F) shape functions against
// Watt2 with n Gauss-Legendre points per axis on [0, 1].
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/cdft.h"

namespace cdft {

struct Pot {
  int kind;  // enum cdft_potential
  double sigma, eps, rcut, alpha;
  double shift, rmin, vmin, r0, r_att_min;
  int bh;

  // the underlying potential from r^2 (vr2 of each class)
  __host__ __device__ double vr2(double r2) const {
    if (kind == CDFT_POT_TWF) {  // Potential1.h:245-254
      if (r2 < sigma * sigma) return 1e50;
      const double y = 1.0 / (r2 / (sigma * sigma) - 1), y3 = y * y * y;
      return (4 * eps / (alpha * alpha)) * (y3 * y3 - alpha * y3);
    }
    if (kind == CDFT_POT_WHDF) {  // Potential1.h:304-309
      const double y2 = sigma * sigma / r2, z2 = rcut * rcut / r2;
      return eps * (y2 - 1) * (z2 - 1) * (z2 - 1);
    }
    const double y2 = sigma * sigma / r2, y6 = y2 * y2 * y2;  // LJ, Potential1.h:176-181
    return 4 * eps * (y6 * y6 - y6);
  }
  __host__ __device__ double vr(double r) const { return vr2(r * r); }
  __host__ __device__ double V(double r) const { return vr(r) - shift; }
  __host__ __device__ double V2(double r2) const { return vr2(r2) - shift; }
  // Potential1::Watt2, Potential1.h:67-75
  __host__ __device__ double watt2(double r2) const {
    if (!(r2 < rcut * rcut)) return 0.0;
    if (bh) return r2 < r0 * r0 ? 0.0 : V2(r2);
    if (r2 < r_att_min * r_att_min) return 0.0;
    if (r2 < rmin * rmin) return vmin;
    return V2(r2);
  }
  // Potential1::V0, Potential1.h:77-81
  __host__ __device__ double V0(double r) const {
    if (bh) return r < r0 ? V(r) : 0.0;
    return r < rmin ? V(r) - vmin : 0.0;
  }
  double hard_core() const { return kind == CDFT_POT_TWF ? sigma : 0.0; }
};

// The constructors: LJ Potential1.h:140-145, tWF :199-206, WHDF :264-271; setBH :47, set_WCA_limit :46.
inline Pot make_pot(int kind, double sigma, double eps, double rcut, double alpha, int bh, double r_att_min) {
  Pot p{};
  p.kind = kind;
  p.sigma = sigma;
  p.eps = eps;
  p.rcut = rcut;
  p.alpha = alpha > 0 ? alpha : 50.0;
  p.bh = 0;
  p.r_att_min = r_att_min;
  if (kind == CDFT_POT_TWF) {
    p.shift = rcut <= 0.0 ? 0.0 : p.vr(rcut);
    p.rmin = sigma * std::sqrt(1 + std::pow(2.0 / p.alpha, 1.0 / 3.0));
    p.vmin = p.V(p.rmin);
    p.r0 = std::sqrt(1 + std::pow(25 * std::sqrt(1 + p.shift) + 25, -1.0 / 3.0));
  } else if (kind == CDFT_POT_WHDF) {
    p.eps = eps * 2 * std::pow(rcut / sigma, 2) * std::pow(2 * ((rcut / sigma) * (rcut / sigma) - 1) / 3.0, -3.0);
    p.shift = 0.0;
    p.rmin = rcut * std::pow((1.0 + 2.0 * (rcut / sigma) * (rcut / sigma)) / 3.0, -0.5);
    p.vmin = p.V(p.rmin);
    p.r0 = 1.0;
  } else {
    p.shift = rcut < 0 ? 0.0 : p.vr(rcut);
    p.rmin = std::pow(2.0, 1.0 / 6.0) * sigma;
    p.vmin = p.V(p.rmin);
    p.r0 = std::pow(0.5 * std::sqrt(1 + p.shift) + 0.5, -1.0 / 6.0);
  }
  if (bh) {
    p.bh = 1;
    p.r_att_min = p.r0;
  }
  return p;
}

constexpr int kMaxRule = 32;

struct MFRule {
  int scheme;  // enum cdft_mf_scheme
  int n;       // points per axis
  double p[kMaxRule], w[kMaxRule];
};

// Gauss-Legendre points and weights on [0, 1] (Newton iteration on P_n from the Chebyshev-like first guess
// cos(pi (i + 3/4) / (n + 1/2)); converged to machine precision).
inline void gauss_legendre01(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      const double dz = p1 / pp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[n - 1 - i] = 0.5 * (1.0 + z);  // ascending on [0, 1]
    w[n - 1 - i] = 1.0 / ((1.0 - z * z) * pp * pp);
  }
}

inline MFRule make_rule(int scheme, int ngauss) {
  MFRule r{};
  r.scheme = scheme;
  auto set = [&](std::initializer_list<double> pts, std::initializer_list<double> wts, double div) {
    r.n = 0;
    for (double v : pts) r.p[r.n++] = v;
    int k = 0;
    for (double v : wts) r.w[k++] = v / div;
  };
  switch (scheme) {
    case CDFT_MF_ZERO: set({0.0}, {1.0}, 1.0); break;  // Interaction.h:301
    case CDFT_MF_LE: set({-2, -1, 0, 1, 2}, {1, 26, 66, 26, 1}, 120.0); break;  // Interaction.h:330-334
    case CDFT_MF_QE:  // Interaction.h:361-364
      set({-2, -1.5, -1, -0.5, 0, 0.5, 1, 1.5, 2}, {-1, 8, 18, 112, 86, 112, 18, 8, -1}, 360.0);
      break;
    case CDFT_MF_LF: set({-1, 0, 1}, {1, 4, 1}, 6.0); break;  // Interaction.h:391-394
    case CDFT_MF_GAUSS_E:
    case CDFT_MF_GAUSS_F:
      r.n = ngauss;
      gauss_legendre01(ngauss, r.p, r.w);
      break;
    default: set({-0.5, 0.0, 0.5}, {1, 1, 1}, 3.0); break;  // QF, Interaction.h:423-426
  }
  return r;
}

// Interaction_Interpolation::generateWeight / Interaction_Gauss::generateWeight for one cell offset S.
__host__ __device__ inline double mf_cell_weight(const MFRule& R, const Pot& v, int Sx, int Sy, int Sz, double dx,
                                                 double dy, double dz) {
  double sum = 0.0;
  if (R.scheme == CDFT_MF_GAUSS_F || R.scheme == CDFT_MF_GAUSS_E) {
    for (int gx = 0; gx < R.n; ++gx)
      for (int gy = 0; gy < R.n; ++gy)
        for (int gz = 0; gz < R.n; ++gz) {
          const double x = R.p[gx], y = R.p[gy], z = R.p[gz];
          double k = 0.0;
          if (R.scheme == CDFT_MF_GAUSS_F) {  // src/Interaction.cpp:465-480
            for (int a = 0; a <= 1; ++a)
              for (int b = 0; b <= 1; ++b)
                for (int c = 0; c <= 1; ++c) {
                  const double ax = Sx + a - 1 + x, by = Sy + b - 1 + y, cz = Sz + c - 1 + z;
                  const double r2 = ax * ax * dx * dx + by * by * dy * dy + cz * cz * dz * dz;
                  k += (a == 0 ? x : 1 - x) * (b == 0 ? y : 1 - y) * (c == 0 ? z : 1 - z) * v.watt2(r2);
                }
          } else {  // src/Interaction.cpp:482-520
            auto f = [](int a, double t) {
              if (a == -2) return t * t * t;
              if (a == -1) return 1 + 3 * t * (1 + t * (1 - t));
              if (a == 0) return 4 + 3 * t * t * (t - 2);
              return (1 - t) * (1 - t) * (1 - t);
            };
            for (int a = -2; a <= 1; ++a)
              for (int b = -2; b <= 1; ++b)
                for (int c = -2; c <= 1; ++c) {
                  const double ax = Sx + a + x, by = Sy + b + y, cz = Sz + c + z;
                  const double r2 = ax * ax * dx * dx + by * by * dy * dy + cz * cz * dz * dz;
                  k += f(a, x) * f(b, y) * f(c, z) * v.watt2(r2);
                }
            k /= 216;
          }
          sum += R.w[gx] * R.w[gy] * R.w[gz] * k;
        }
    return sum;
  }
  for (int i = 0; i < R.n; ++i)
    for (int j = 0; j < R.n; ++j)
      for (int k = 0; k < R.n; ++k) {
        const double ax = Sx + R.p[i], by = Sy + R.p[j], cz = Sz + R.p[k];
        const double r2 = ax * ax * dx * dx + by * by * dy * dy + cz * cz * dz * dz;
        sum += R.w[i] * R.w[j] * R.w[k] * v.watt2(r2);
      }
  return sum;
}

}  // namespace cdft
