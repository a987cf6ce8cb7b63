// Pointwise FMT free-energy density Phi(n_alpha) and its derivatives, in FP64,
// for the five reference models, collapsed to the 11 stored dPhi arrays.
//
// Reference: FMT::calculate_Phi / calculate_dPhi_wrt_fundamental_measures
// (src/FMT.cpp:114-176), the model hooks in include/FMT.h:139-643, the measure
// construction FMT::getWeightedDensities (src/FMT.cpp:78-111) and the collapse
// FMT_Species::set_fundamental_measure_derivatives (include/Species.h:310-329).
//
// Array order (include/Species.h:351-360): 0 eta, 1 s, 2..4 v, 5 Txx, 6 Txy,
// 7 Txz, 8 Tyy, 9 Tyz, 10 Tzz.  The tensor is symmetric by construction, so
// T.T, v.T and T.v are evaluated from the six independent entries.
#pragma once
#include <hip/hip_runtime.h>

namespace cdft {

enum { M_ROSENFELD = 0, M_RSLT = 1, M_ESFMT = 2, M_WBI = 3, M_WBII = 4 };

constexpr double kPi = 3.14159265358979323846;

struct PhiParams {
  double hsd;
  double A, B;  // esFMT coefficients (WB I/II: 1, -1)
  double inv_d, inv_d2;  // 1/hsd, 1/hsd^2
  double ca, cb, ea, eb;  // A/(24 pi), B/(24 pi), A/(8 pi), B/(4 pi): site-independent factors of the tensor terms
};

// WB I and WB II are esFMT(1, -1) whatever A and B the caller holds (FMT.h:505-643).
inline PhiParams make_phi_params(double hsd, double A, double B, bool whitebear = false) {
  if (whitebear) {
    A = 1.0;
    B = -1.0;
  }
  return PhiParams{hsd, A, B, 1.0 / hsd, 1.0 / (hsd * hsd), A / (24 * kPi), B / (24 * kPi), A / (8 * kPi),
                   B / (4 * kPi)};
}

// Forward-mode dual number v + d e (e^2 = 0): evaluating the derivative code on
// n + psi e gives, in the e part, the second derivative of Phi applied to psi
// (the Hessian-vector product of FMT::calculate_d2Phi_dot_V, src/FMT.cpp:181-280).
struct Dual {
  double v, d;
  __host__ __device__ constexpr Dual(double a = 0.0, double b = 0.0) : v(a), d(b) {}
};
__host__ __device__ inline Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual operator+(const Dual& a, double b) { return Dual(a.v + b, a.d); }
__host__ __device__ inline Dual operator+(double a, const Dual& b) { return Dual(a + b.v, b.d); }
__host__ __device__ inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
__host__ __device__ inline Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual operator-(const Dual& a, double b) { return Dual(a.v - b, a.d); }
__host__ __device__ inline Dual operator-(double a, const Dual& b) { return Dual(a - b.v, -b.d); }
__host__ __device__ inline Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline Dual operator*(const Dual& a, double b) { return Dual(a.v * b, a.d * b); }
__host__ __device__ inline Dual operator*(double a, const Dual& b) { return Dual(a * b.v, a * b.d); }
__host__ __device__ inline Dual operator/(const Dual& a, double b) { return Dual(a.v / b, a.d / b); }
__host__ __device__ inline Dual operator/(double a, const Dual& b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
__host__ __device__ inline Dual operator/(const Dual& a, const Dual& b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline bool operator<(const Dual& a, double b) { return a.v < b; }
__host__ __device__ inline double flog(double x) { return log(x); }
__host__ __device__ inline Dual flog(const Dual& x) { return Dual(log(x.v), x.d / x.v); }

// eta-dependent cofactors f1, f2, f3 and first derivatives.  Written with the
// two reciprocals 1/(1-x) and 1/x instead of one division per term (a double
// division is ~10 instructions on the GPU); same formulas as the reference.
// Both reciprocals are correctly rounded, as the reference's divisions are:
// f3 and f3' cancel terms of order 1/x^2 and 1/x^3 at small x, so a reciprocal
// one ulp off would move them far beyond the rounding of the reference.
template <int MODEL, typename S = double>
__host__ __device__ inline void fcoef(S x, S& f1, S& f1p, S& f2, S& f2p, S& f3, S& f3p) {
  const S om = 1.0 - x;
  const bool series = x < 1e-12;
  const S f = 1.0 / om;
  f1 = flog(om);  // FMT.h:153-156
  f1p = -f;
  if constexpr (MODEL == M_WBII) {  // FMT.h:588-632
    if (series) {
      f2 = 1 + x * (1 + x * ((10.0 / 9) + (7.0 / 6) * x));
      f2p = 1 + x * ((20.0 / 9) + 3.5 * x);
      f3 = 1.5 + (7.0 / 3) * x + 3.25 * x * x + 4.2 * x * x * x;
      f3p = (7.0 / 3) + 7.5 * x + 12.6 * x * x;
    } else {
      const S L = f1, ix = 1.0 / x, ix2 = ix * ix, f2_ = f * f;
      f2 = (1.0 / 3.0) + (4.0 / 3.0) * f + (2.0 / 3.0) * ix * L;
      f2p = (2.0 / 3.0) * (3 * x - 1) * ix * f2_ - (2.0 / 3.0) * ix2 * L;
      f3 = -(1.0 - 3 * x + x * x) * ix * f2_ - ix2 * L;
      f3p = (2 - 5 * x + 6 * x * x - x * x * x) * ix2 * f2_ * f + 2.0 * ix2 * ix * L;
    }
  } else {
    f2 = f;  // FMT.h:168-183
    f2p = f * f;
    if constexpr (MODEL == M_RSLT || MODEL == M_WBI) {  // FMT.h:267-281, 536-550
      if (series) {
        f3 = 1.5 + (8.0 / 3) * x + 3.75 * x * x + 4.8 * x * x * x;
        f3p = (8.0 / 3) + 7.5 * x + 14.4 * x * x;
      } else {
        const S ix = 1.0 / x, ix2 = ix * ix, f2_ = f * f;
        f3 = ix * f2_ + f1 * ix2;
        f3p = (-2 + 5 * x - x * x) * ix2 * f2_ * f - 2 * ix2 * ix * f1;
      }
    } else {  // Rosenfeld, esFMT: FMT.h:185-195
      f3 = f * f;
      f3p = 2 * f * f * f;
    }
  }
}

// Derivatives of Phi with respect to the measures: eta, s0, s1, s2, v1[3], v2[3] and the six independent tensor
// entries (T = xx, xy, xz, yy, yz, zz, each the derivative with respect to one of T_jk, T_kj).
template <typename S>
struct PhiDeriv {
  S eta, s0, s1, s2, v1[3], v2[3], T[6];
};

// Phi and its derivatives from the fundamental measures (FMT::calculate_Phi and
// calculate_dPhi_wrt_fundamental_measures, src/FMT.cpp:114-176; model hooks include/FMT.h:139-643).
template <int MODEL, typename S = double>
__host__ __device__ inline S phi_meas(const S& eta, const S& s0, const S& s1, const S& s2, const S (&v1)[3],
                                      const S (&v2)[3], const S* T6, const PhiParams& pp, PhiDeriv<S>& d) {
  const S v1v2 = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
  const S v2v2 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];

  S f1, f1p, f2, f2p, f3, f3p;
  fcoef<MODEL, S>(eta, f1, f1p, f2, f2p, f3, f3p);

  S P3, dS2, dV2[3];
  S dT[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // xx, xy, xz, yy, yz, zz
  if constexpr (MODEL == M_ROSENFELD) {  // FMT.h:202-240
    P3 = (1.0 / (24 * kPi)) * (s2 * s2 * s2 - 3 * s2 * v2v2);
    dS2 = (1.0 / (24 * kPi)) * (3 * s2 * s2 - 3 * v2v2);
    for (int k = 0; k < 3; ++k) dV2[k] = (1.0 / (24 * kPi)) * (-6 * s2 * v2[k]);
  } else if constexpr (MODEL == M_RSLT) {  // FMT.h:293-358
    if (s2 < 1e-20) {
      P3 = 0.0;
      dS2 = 0.0;
      dV2[0] = dV2[1] = dV2[2] = 0.0;
    } else {
      const S psi = v2v2 / (s2 * s2);
      const S q = 1.0 - psi;
      P3 = (1.0 / (36 * kPi)) * s2 * s2 * s2 * q * q * q;
      dS2 = (1.0 / (12 * kPi)) * s2 * s2 * (1 + psi) * q * q;
      for (int k = 0; k < 3; ++k) dV2[k] = -(1.0 / (6 * kPi)) * s2 * v2[k] * q * q;
    }
  } else {  // esFMT family: FMT.h:396-455
    const S Txx = T6[0], Txy = T6[1], Txz = T6[2], Tyy = T6[3], Tyz = T6[4], Tzz = T6[5];
    // TT = T.T (symmetric)
    const S Qxx = Txx * Txx + Txy * Txy + Txz * Txz;
    const S Qxy = Txx * Txy + Txy * Tyy + Txz * Tyz;
    const S Qxz = Txx * Txz + Txy * Tyz + Txz * Tzz;
    const S Qyy = Txy * Txy + Tyy * Tyy + Tyz * Tyz;
    const S Qyz = Txy * Txz + Tyy * Tyz + Tyz * Tzz;
    const S Qzz = Txz * Txz + Tyz * Tyz + Tzz * Tzz;
    const S vT0 = v2[0] * Txx + v2[1] * Txy + v2[2] * Txz;
    const S vT1 = v2[0] * Txy + v2[1] * Tyy + v2[2] * Tyz;
    const S vT2 = v2[0] * Txz + v2[1] * Tyz + v2[2] * Tzz;
    const S vTv = v2[0] * vT0 + v2[1] * vT1 + v2[2] * vT2;
    const S T2 = Qxx + Qyy + Qzz;
    const S T3 = Qxx * Txx + Qyy * Tyy + Qzz * Tzz + 2 * (Qxy * Txy + Qxz * Txz + Qyz * Tyz);
    const double ca = pp.ca, cb = pp.cb;
    P3 = ca * (s2 * s2 * s2 - 3 * s2 * v2v2 + 3 * vTv - T3) + cb * (s2 * s2 * s2 - 3 * s2 * T2 + 2 * T3);
    dS2 = ca * (3 * s2 * s2 - 3 * v2v2) + cb * (3 * s2 * s2 - 3 * T2);
    dV2[0] = ca * (-6 * s2 * v2[0] + 6 * vT0);
    dV2[1] = ca * (-6 * s2 * v2[1] + 6 * vT1);
    dV2[2] = ca * (-6 * s2 * v2[2] + 6 * vT2);
    const double ea = pp.ea, eb = pp.eb;
    dT[0] = ea * (v2[0] * v2[0] - Qxx) + eb * (-s2 * Txx + Qxx);
    dT[1] = ea * (v2[0] * v2[1] - Qxy) + eb * (-s2 * Txy + Qxy);
    dT[2] = ea * (v2[0] * v2[2] - Qxz) + eb * (-s2 * Txz + Qxz);
    dT[3] = ea * (v2[1] * v2[1] - Qyy) + eb * (-s2 * Tyy + Qyy);
    dT[4] = ea * (v2[1] * v2[2] - Qyz) + eb * (-s2 * Tyz + Qyz);
    dT[5] = ea * (v2[2] * v2[2] - Qzz) + eb * (-s2 * Tzz + Qzz);
  }

  const double c1 = 1.0 / kPi, c2 = 1.0 / (2 * kPi);
  const S phi = -c1 * s0 * f1 + c2 * (s1 * s2 - v1v2) * f2 + P3 * f3;
  d.eta = -c1 * s0 * f1p + c2 * (s1 * s2 - v1v2) * f2p + P3 * f3p;
  d.s0 = -c1 * f1;
  d.s1 = c2 * s2 * f2;
  d.s2 = c2 * s1 * f2 + dS2 * f3;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    d.v1[k] = -c2 * v2[k] * f2;
    d.v2[k] = -c2 * v1[k] * f2 + dV2[k] * f3;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) d.T[k] = dT[k] * f3;
  return phi;
}

// Collapse onto the 11 stored arrays of a species of diameter 1/id (FMT_Species::set_fundamental_measure_derivatives,
// include/Species.h:310-329): eta; s0/d^2 + s1/d + s2; the vector with the parity sign; tensor off-diagonals twice.
template <bool TENSOR, typename S>
__host__ __device__ inline void phi_collapse(const PhiDeriv<S>& d, double id, double id2, S* out) {
  out[0] = d.eta;
  out[1] = d.s0 * id2 + d.s1 * id + d.s2;
#pragma unroll
  for (int k = 0; k < 3; ++k) out[2 + k] = -(d.v2[k] + d.v1[k] * id);  // parity sign, Species.h:324
  if constexpr (TENSOR) {
    out[5] = d.T[0];
    out[6] = 2 * d.T[1];  // off-diagonals counted twice, Species.h:327
    out[7] = 2 * d.T[2];
    out[8] = d.T[3];
    out[9] = 2 * d.T[4];
    out[10] = d.T[5];
  } else {
#pragma unroll
    for (int a = 5; a < 11; ++a) out[a] = S(0.0);
  }
}

// Returns Phi at one site; writes the 11 collapsed derivatives to out[].
// n[] holds the 11 weighted densities of the site (tensor entries ignored for
// Rosenfeld / RSLT, whose tensor arrays are never convolved).
// S = double: Phi and its collapsed first derivatives.  S = Dual with n = n_rho + psi e: the e parts of out[]
// are the collapsed second derivatives applied to psi (same collapse and parity sign).
template <int MODEL, typename S = double>
__host__ __device__ inline S phi_site(const S* n, const PhiParams& pp, S* out) {
  constexpr bool TENSOR = MODEL >= M_ESFMT;
  const double id = pp.inv_d, id2 = pp.inv_d2;
  const S eta = n[0];
  const S s0 = n[1] * id2, s1 = n[1] * id, s2 = n[1];
  const S v2[3] = {n[2], n[3], n[4]};
  const S v1[3] = {n[2] * id, n[3] * id, n[4] * id};
  PhiDeriv<S> d;
  const S phi = phi_meas<MODEL, S>(eta, s0, s1, s2, v1, v2, n + 5, pp, d);
  phi_collapse<TENSOR, S>(d, id, id2, out);
  return phi;
}

}  // namespace cdft
