// Cell-integrated FMT weights of one lattice cell, generic in the arithmetic
// type (double, or dd = double-double for the device generator).
//
// The integrals are those of FMT_Species::generateWeights
// (src/FMT_Species.cpp:310-496): the sphere's volume, surface, vector and
// tensor densities against the trilinear hat of the cell, from the G
// functions of :148-308 evaluated at the corners of the 8 adjoining cells.
// Two changes make them well conditioned:
//  * every asin(p/q) whose complementary leg is known exactly is written
//    atan2(p, leg): with E = R^2 - X^2 - V^2 the angles are atan2(X, sqrt E),
//    atan2(V, sqrt E) and atan2(X V, R sqrt E), and at the upper limit
//    X = sqrt(R^2 - Va^2 - Vb^2) one has E_a = Vb^2 and E_b = Va^2 exactly;
//  * the corner coordinates d*k are formed exactly (two_prod) and E, the
//    polynomial parts and the 8-corner sums are carried in double-double.
//    The alternating corner sums cancel G values of order R^6 down to a cell
//    weight of order dx^3, and sqrt(E) amplifies the rounding of E where a
//    corner lies close to the sphere; in plain double both grow as dx
//    shrinks (about 1e-11 dV per cell at dx = 0.04, R = 0.5).
#pragma once
#include "dd.hpp"

namespace cdft {

CDFT_HD inline double r_sqrt(double a) { return std::sqrt(a); }
CDFT_HD inline double r_atan2(double y, double x) { return std::atan2(y, x); }
CDFT_HD inline dd r_sqrt(dd a) { return dd_sqrt(a); }
CDFT_HD inline dd r_atan2(dd y, dd x) { return dd_atan2(y, x); }

template <typename R>
struct RT;
template <>
struct RT<double> {
  CDFT_HD static double pi() { return 3.14159265358979323846; }
  CDFT_HD static double coord(double d, int k) { return d * double(k); }
};
template <>
struct RT<dd> {
  CDFT_HD static dd pi() { return dd_pi(); }
  CDFT_HD static dd coord(double d, int k) { return two_prod(d, double(k)); }
};

// I and J moment integrals of one (X, V) pair (getI / getJ, src/FMT_Species.cpp:112-145) in the atan2 form.
template <typename R>
CDFT_HD void wc_moments(R X, R V, R Rr, R E, R* I, R* J) {
  CDFT_EXACT_FP
  if (E < 0.0) E = R(0.0);
  const R sE = r_sqrt(E);
  const R aX = r_atan2(X, sE);
  const R aV = r_atan2(V, sE);
  const R aVX = r_atan2(X * V, Rr * sE);
  const R A2 = X * X + E;  // R^2 - V^2
  const R X2 = X * X, R2 = Rr * Rr, V2 = V * V;
  const R c3 = R(1.0) / 3.0, c6 = R(1.0) / 6.0, c15 = R(1.0) / 15.0, c48 = R(1.0) / 48.0, c40 = R(1.0) / 40.0,
          c5 = R(1.0) / 5.0, c12 = R(1.0) / 12.0;
  I[0] = 0.5 * X * sE + 0.5 * A2 * aX;
  I[1] = -(c3 * E * sE);
  I[2] = 0.125 * X * (2.0 * X2 - A2) * sE + 0.125 * A2 * A2 * aX;
  I[3] = -(c15 * E * (3.0 * X2 + 2.0 * A2) * sE);
  I[4] = c48 * (8.0 * X2 * X2 - 2.0 * A2 * X2 - 3.0 * A2 * A2) * X * sE + 0.0625 * A2 * A2 * A2 * aX;
  J[0] = X * aV + V * aX - Rr * aVX;
  J[1] = -(0.5 * V * sE) + 0.5 * (X2 - R2) * aV;
  J[2] = -(c6 * V * (V2 - 3.0 * R2) * aX) + c3 * X2 * X * aV - c3 * R2 * Rr * aVX - c6 * V * X * sE;
  J[3] = 0.25 * (X2 * X2 - R2 * R2) * aV + (0.125 * RT<R>::pi()) * R2 * R2 - c12 * V * (5.0 * R2 + X2 - 2.0 * V2) * sE;
  J[4] = c40 * X * V * (3.0 * V2 - 2.0 * X2 - 7.0 * R2) * sE + c40 * V * (3.0 * V2 * V2 - 10.0 * V2 * R2 + 15.0 * R2 * R2) * aX +
         c5 * X2 * X2 * X * aV - c5 * R2 * R2 * Rr * aVX;
}

// G functions, kind 0 eta (:148-193), 1 s (:195-221), 2 v_x (:223-250), 3 T_xx (:252-278), 4 T_xy (:280-308).
// Ey = R^2 - X^2 - Vy^2, Ez = R^2 - X^2 - Vz^2.
template <typename R>
CDFT_HD R wc_G(int kind, R Rr, R X, R Vy, R Vz, R Tx, R Ty, R Tz, R Ey, R Ez) {
  CDFT_EXACT_FP
  R Iy[5], Iz[5], Jy[5], Jz[5];
  wc_moments(X, Vy, Rr, Ey, Iy, Jy);
  wc_moments(X, Vz, Rr, Ez, Iz, Jz);
  const R pi = RT<R>::pi();
  const R R2 = Rr * Rr, X2 = X * X;
  const R c3 = R(1.0) / 3.0, c6 = R(1.0) / 6.0, c48 = R(1.0) / 48.0, c40 = R(1.0) / 40.0, c10 = R(1.0) / 10.0,
          c12 = R(1.0) / 12.0, c5 = R(1.0) / 5.0;
  if (kind == 0) {
    const R A = Ty * Tz * Vy * Vz + 0.25 * Vy * Vy * Vz * Vz + 0.125 * R2 * R2 -
                c6 * (Ty * Vy * Vy * Vy + Tz * Vz * Vz * Vz) - 0.5 * Tz * Vz * Vy * Vy - 0.5 * Ty * Vy * Vz * Vz +
                0.125 * Vy * Vy * Vy * Vy + 0.125 * Vz * Vz * Vz * Vz +
                0.5 * R2 * (Ty * Vy + Tz * Vz - 0.5 * Vy * Vy - 0.5 * Vz * Vz + 0.5 * pi * Ty * Tz);
    const R B = 0.5 * Ty * Vy + 0.5 * Tz * Vz - 0.25 * Vy * Vy - 0.25 * Vz * Vz + 0.25 * pi * Ty * Tz + 0.25 * R2;
    const R C = 0.5 * Ty * Vy + c3 * (R2 - Vy * Vy);
    const R D = 0.5 * Tz * Vz + c3 * (R2 - Vz * Vz);
    R g = Tx * A * X - 0.5 * A * X2 - c3 * B * Tx * X2 * X + 0.25 * X2 * X2 * B + c40 * Tx * X2 * X2 * X -
          c48 * X2 * X2 * X2;
    g -= 0.5 * Ty * Tz * (Tx * R2 * Jy[0] - R2 * Jy[1] - Tx * Jy[2] + Jy[3]);
    g -= 0.5 * Ty * Tz * (Tx * R2 * Jz[0] - R2 * Jz[1] - Tx * Jz[2] + Jz[3]);
    g -= Tx * Tz * C * Iy[0] - Tz * C * Iy[1] - c3 * Tx * Tz * Iy[2] + c3 * Tz * Iy[3];
    g -= Tx * Ty * D * Iz[0] - Ty * D * Iz[1] - c3 * Tx * Ty * Iz[2] + c3 * Ty * Iz[3];
    return g;
  }
  if (kind == 4) {
    const R A = c6 * (3.0 * Ty * R2 + 2.0 * Vy * Vy * Vy - 3.0 * Ty * Vy * Vy - 3.0 * Ty * Vz * Vz + 6.0 * Ty * Tz * Vz +
                      1.5 * pi * R2 * Tz);
    const R B = 0.25 * pi * Tz + 0.5 * Ty;
    const R C = -(c6 * (2.0 * R2 - 2.0 * Vz * Vz + 3.0 * Tz * Vz));
    R g = -(0.5 * Rr * A * Tx * X2) + c3 * Rr * A * X2 * X + 0.25 * Rr * B * Tx * X2 * X2 - c5 * Rr * B * X2 * X2 * X;
    g += 0.5 * Rr * Tz * (Tx * R2 * Jy[1] - R2 * Jy[2] - Tx * Jy[3] + Jy[4]);
    g += 0.5 * Rr * Tz * (Tx * R2 * Jz[1] - R2 * Jz[2] - Tx * Jz[3] + Jz[4]);
    g += 0.5 * Rr * Tz * (2.0 * Ty - Vy) * (Tx * Iy[1] - Iy[2]);
    g += Rr * (C * Iz[2] - Tx * C * Iz[1] - c3 * Tx * Iz[3] + c3 * Iz[4]);
    return g;
  }
  const int o = kind - 1;  // moment order of s, v_x, T_xx
  const R A = Ty * Vy + Tz * Vz - 0.5 * Vy * Vy - 0.5 * Vz * Vz + 0.5 * R2 + 0.5 * pi * Ty * Tz;
  R g;
  if (o == 0)
    g = 0.125 * Rr * X2 * X2 - c6 * Rr * Tx * X2 * X - 0.5 * Rr * A * X2 + Rr * A * Tx * X;
  else if (o == 1)
    g = c10 * Rr * X2 * X2 * X - 0.125 * Rr * Tx * X2 * X2 - c3 * Rr * A * X2 * X + 0.5 * Rr * A * Tx * X2;
  else
    g = c12 * Rr * X2 * X2 * X2 - c10 * Rr * Tx * X2 * X2 * X - 0.25 * Rr * A * X2 * X2 + c3 * Rr * A * Tx * X2 * X;
  g -= Rr * Tz * (Tx * Iy[o] - Iy[o + 1]);
  g -= Rr * Ty * (Tx * Iz[o] - Iz[o + 1]);
  g -= Rr * Ty * Tz * (Tx * Jy[o] - Jy[o + 1]);
  g -= Rr * Ty * Tz * (Tx * Jz[o] - Jz[o + 1]);
  return g;
}

// Roles of the 11 weights (include/Species.h:351-360 order): G kind, the axis playing X and the two others
// (a, b), the axes whose parity signs multiply the weight, and the power of 1/R.
struct WRole {
  int kind, ax, aa, ab, sgn_mask, rpow;
};
CDFT_HD inline WRole weight_role(int a) {
  CDFT_EXACT_FP
  switch (a) {
    case 0: return {0, 0, 1, 2, 0, 0};  // eta
    case 1: return {1, 0, 1, 2, 0, 0};  // s
    case 2: return {2, 0, 1, 2, 1, 1};  // v_x
    case 3: return {2, 1, 0, 2, 2, 1};  // v_y
    case 4: return {2, 2, 1, 0, 4, 1};  // v_z
    case 5: return {3, 0, 1, 2, 0, 2};  // T_xx
    case 6: return {4, 0, 1, 2, 3, 2};  // T_xy
    case 7: return {4, 0, 2, 1, 5, 2};  // T_xz
    case 8: return {3, 1, 0, 2, 0, 2};  // T_yy
    case 9: return {4, 1, 2, 0, 6, 2};  // T_yz
    default: return {3, 2, 1, 0, 0, 2};  // T_zz
  }
}

// Weight a of the cell (Sx, Sy, Sz) of the positive octant (before the octant scatter), sphere radius Rd.
template <typename R>
CDFT_HD double wc_cell(int a, int Sx, int Sy, int Sz, double dx, double dy, double dz, double Rd) {
  CDFT_EXACT_FP
  const double d[3] = {dx, dy, dz};
  const int S[3] = {Sx, Sy, Sz};
  const R Rr(Rd), R2 = Rr * Rr;
  const R dV = R(dx) * dy * dz;
  R r2min(0.0), r2max(0.0);
  for (int i = 0; i < 3; ++i) {
    const R lo = RT<R>::coord(d[i], S[i] == 0 ? 0 : S[i] - 1), hi = RT<R>::coord(d[i], S[i] + 1);
    r2min += lo * lo;
    r2max += hi * hi;
  }
  if (R2 > r2max) return a == 0 ? to_double(dV) : 0.0;
  if (!(R2 > r2min)) return 0.0;
  const WRole ro = weight_role(a);
  R scale = R(1.0) / dV;
  for (int k = 0; k < ro.rpow; ++k) scale = scale / Rr;
  R w(0.0);
  for (int oct = 0; oct < 8; ++oct) {
    int kT[3], step[3], par[3];  // T_i = d_i kT_i, corner offsets step_i in {0, kT_i - kV_i}
    for (int i = 0; i < 3; ++i) {
      const int e = ((oct >> (2 - i)) & 1) ? 1 : -1;
      int st = e, kt = S[i] + e;
      par[i] = 1;
      if (kt < 0) {  // the octant image across the axis: T = d, parity -1 (src/FMT_Species.cpp:392-406)
        par[i] = -1;
        st = 1;
        kt = 1;
      }
      kT[i] = kt;
      step[i] = st;
    }
    int sg = 1;
    for (int i = 0; i < 3; ++i)
      if ((ro.sgn_mask >> i) & 1) sg *= par[i];
    R T[3];
    for (int i = 0; i < 3; ++i) T[i] = RT<R>::coord(d[i], kT[i]);
    for (int j = 0; j < 8; ++j) {
      int kV[3];
      for (int i = 0; i < 3; ++i) kV[i] = kT[i] - (((j >> (2 - i)) & 1) ? step[i] : 0);
      R V[3];
      for (int i = 0; i < 3; ++i) V[i] = RT<R>::coord(d[i], kV[i]);
      const R E0 = R2 - V[0] * V[0] - V[1] * V[1] - V[2] * V[2];
      if (E0 <= 0.0) continue;  // corner outside the sphere
      const R X = V[ro.ax], Va = V[ro.aa], Vb = V[ro.ab];
      const R Tx = T[ro.ax], Ta = T[ro.aa], Tb = T[ro.ab];
      const R h = r_sqrt(R2 - Va * Va - Vb * Vb);
      const R glo = wc_G<R>(ro.kind, Rr, X, Va, Vb, Tx, Ta, Tb, R2 - X * X - Va * Va, R2 - X * X - Vb * Vb);
      const R ghi = wc_G<R>(ro.kind, Rr, h, Va, Vb, Tx, Ta, Tb, Vb * Vb, Va * Va);
      const int cs = (__builtin_popcount(j) & 1) ? -sg : sg;
      if (cs > 0)
        w += glo - ghi;
      else
        w -= glo - ghi;
    }
  }
  return to_double(w * scale);
}

}  // namespace cdft
