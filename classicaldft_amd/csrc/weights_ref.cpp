// FMT cell weights by the reference's own expressions, in double on the host
// (cdft_system_desc::weights_mode = CDFT_WEIGHTS_REFERENCE).
//
// FMT_Species::generateWeights (src/FMT_Species.cpp:310-496) with getI / getJ
// (:112-145) and the G functions (:148-308): the asin forms with their clamps
// and 1e-15 guards, the R += 1e-15 shift inside every G, the same order of
// every floating-point operation, the same corner test and the same scatter
// order onto the periodic grid.  Built without floating-point contraction, so
// on an x86-64 host with glibc's libm the result is the one a reference build
// without FMA contraction produces.  These weights carry the reference's
// rounding noise (asin near +-1 turns 1e-16 into 1e-8); the device generator
// (wcells.hpp) computes the same integrals without it.
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "weights_ref.hpp"

namespace cdft {
namespace {

// src/FMT_Species.cpp:112-122
void ref_I(double X, double A, double I[]) {
  const double a = std::asin(X / A);
  const double b = std::sqrt(std::fabs(A * A - X * X));
  I[0] = 0.5 * X * b + 0.5 * A * A * a;
  I[1] = -(1.0 / 3) * (A * A - X * X) * b;
  I[2] = 0.125 * X * (2 * X * X - A * A) * b + 0.125 * A * A * A * A * a;
  I[3] = -(1.0 / 15) * (A * A - X * X) * (3 * X * X + 2 * A * A) * b;
  I[4] = (1.0 / 48) * (8 * X * X * X * X - 2 * A * A * X * X - 3 * A * A * A * A) * X * b + 0.0625 * A * A * A * A * A * A * a;
}

// src/FMT_Species.cpp:124-145
void ref_J(double X, double V, double R, double J[]) {
  const double small = 1e-15;
  const double aV = std::asin(std::max(-1.0, std::min(1.0, V / std::max(small, std::sqrt(R * R - X * X)))));
  const double aX = std::asin(std::max(-1.0, std::min(1.0, X / std::sqrt(R * R - V * V))));
  const double aVX =
      std::asin(std::max(-1.0, std::min(1.0, X * V / std::max(small, std::sqrt((R * R - V * V) * (R * R - X * X))))));
  const double b = std::sqrt(std::fabs(R * R - V * V - X * X));
  J[0] = X * aV + V * aX - R * aVX;
  J[1] = -0.5 * V * b + 0.5 * (X * X - R * R) * aV;
  J[2] = -(1.0 / 6) * V * (V * V - 3 * R * R) * aX + (1.0 / 3) * X * X * X * aV - (1.0 / 3) * R * R * R * aVX -
         (1.0 / 6) * V * X * b;
  J[3] = 0.25 * (X * X * X * X - R * R * R * R) * aV + (M_PI / 8) * R * R * R * R - (V / 12.0) * (5 * R * R + X * X - 2 * V * V) * b;
  J[4] = 0.025 * X * V * (3 * V * V - 2 * X * X - 7 * R * R) * b +
         0.025 * V * (3 * V * V * V * V - 10 * V * V * R * R + 15 * R * R * R * R) * aX + 0.2 * X * X * X * X * X * aV -
         0.2 * R * R * R * R * R * aVX;
}

enum { K_ETA, K_S, K_VX, K_TXX, K_TXY };

// G_eta, G_s, G_vx, G_txx, G_txy (src/FMT_Species.cpp:148-308)
double ref_G(int kind, double R, double X, double Vy, double Vz, double Tx, double Ty, double Tz) {
  double Iy[5], Iz[5], Jy[5], Jz[5];
  R += 1e-15;
  ref_I(X, std::sqrt(R * R - Vy * Vy), Iy);
  ref_I(X, std::sqrt(R * R - Vz * Vz), Iz);
  ref_J(X, Vy, R, Jy);
  ref_J(X, Vz, R, Jz);
  double g;
  switch (kind) {
    case K_ETA: {
      const double A = Ty * Tz * Vy * Vz + 0.25 * Vy * Vy * Vz * Vz + 0.125 * R * R * R * R -
                       (1.0 / 6) * (Ty * Vy * Vy * Vy + Tz * Vz * Vz * Vz) - 0.5 * Tz * Vz * Vy * Vy -
                       0.5 * Ty * Vy * Vz * Vz + 0.125 * Vy * Vy * Vy * Vy + 0.125 * Vz * Vz * Vz * Vz +
                       0.5 * R * R * (Ty * Vy + Tz * Vz - 0.5 * Vy * Vy - 0.5 * Vz * Vz + 0.5 * M_PI * Ty * Tz);
      const double B = 0.5 * Ty * Vy + 0.5 * Tz * Vz - 0.25 * Vy * Vy - 0.25 * Vz * Vz + (M_PI / 4) * Ty * Tz + 0.25 * R * R;
      const double C = 0.5 * Ty * Vy + (1.0 / 3) * (R * R - Vy * Vy);
      const double D = 0.5 * Tz * Vz + (1.0 / 3) * (R * R - Vz * Vz);
      g = Tx * A * X - 0.5 * A * X * X - (1.0 / 3) * B * Tx * X * X * X + 0.25 * X * X * X * X * B +
          0.025 * Tx * X * X * X * X * X - (1.0 / 48) * X * X * X * X * X * X;
      g -= 0.5 * Ty * Tz * (Tx * R * R * Jy[0] - R * R * Jy[1] - Tx * Jy[2] + Jy[3]);
      g -= 0.5 * Ty * Tz * (Tx * R * R * Jz[0] - R * R * Jz[1] - Tx * Jz[2] + Jz[3]);
      g -= (Tx * Tz * C * Iy[0] - Tz * C * Iy[1] - (1.0 / 3) * Tx * Tz * Iy[2] + (1.0 / 3) * Tz * Iy[3]);
      g -= (Tx * Ty * D * Iz[0] - Ty * D * Iz[1] - (1.0 / 3) * Tx * Ty * Iz[2] + (1.0 / 3) * Ty * Iz[3]);
      return g;
    }
    case K_TXY: {
      const double A = (1.0 / 6) * (3 * Ty * R * R + 2 * Vy * Vy * Vy - 3 * Ty * Vy * Vy - 3 * Ty * Vz * Vz +
                                    6 * Ty * Tz * Vz + 1.5 * M_PI * R * R * Tz);
      const double B = 0.25 * M_PI * Tz + 0.5 * Ty;
      const double C = -(1.0 / 6) * (2 * R * R - 2 * Vz * Vz + 3 * Tz * Vz);
      g = -0.5 * R * A * Tx * X * X + (1.0 / 3) * R * A * X * X * X + 0.25 * R * B * Tx * X * X * X * X -
          0.2 * R * B * X * X * X * X * X;
      g += 0.5 * R * Tz * (Tx * R * R * Jy[1] - R * R * Jy[2] - Tx * Jy[3] + Jy[4]);
      g += 0.5 * R * Tz * (Tx * R * R * Jz[1] - R * R * Jz[2] - Tx * Jz[3] + Jz[4]);
      g += 0.5 * R * Tz * (2 * Ty - Vy) * (Tx * Iy[1] - Iy[2]);
      g += R * (-Tx * C * Iz[1] + C * Iz[2] - (1.0 / 3) * Tx * Iz[3] + (1.0 / 3) * Iz[4]);
      return g;
    }
    default: {
      // G_s, G_vx and G_txx differ in their polynomial and in the moment order o
      const double A = Ty * Vy + Tz * Vz - 0.5 * Vy * Vy - 0.5 * Vz * Vz + 0.5 * R * R + (M_PI / 2) * Ty * Tz;
      const int o = kind - K_S;
      if (kind == K_S)
        g = 0.125 * R * X * X * X * X - (1.0 / 6) * R * Tx * X * X * X - 0.5 * R * A * X * X + R * A * Tx * X;
      else if (kind == K_VX)
        g = 0.1 * R * X * X * X * X * X - 0.125 * R * Tx * X * X * X * X - (1.0 / 3) * R * A * X * X * X +
            0.5 * R * A * Tx * X * X;
      else
        g = (1.0 / 12) * R * X * X * X * X * X * X - 0.1 * R * Tx * X * X * X * X * X - 0.25 * R * A * X * X * X * X +
            (1.0 / 3) * R * A * Tx * X * X * X;
      g -= R * Tz * (Tx * Iy[o] - Iy[o + 1]);
      g -= R * Ty * (Tx * Iz[o] - Iz[o + 1]);
      g -= R * Ty * Tz * (Tx * Jy[o] - Jy[o + 1]);
      g -= R * Ty * Tz * (Tx * Jz[o] - Jz[o + 1]);
      return g;
    }
  }
}

}  // namespace

void reference_weights(int Nx, int Ny, int Nz, double dx, double dy, double dz, double hsd, double* w) {
  const long N = (long)Nx * Ny * Nz;
  const double dV = dx * dy * dz;
  const double R = hsd / 2;
  if (hsd < std::min(std::min(dx, dy), dz)) throw std::runtime_error("hsd is less than the lattice spacing");
  const int Sx_max = 2 + int(R / dx), Sy_max = 2 + int(R / dy), Sz_max = 2 + int(R / dz);
  std::fill(w, w + 11 * N, 0.0);
  const int I[2] = {-1, 1};
  for (int Sx = 0; Sx <= Sx_max; Sx++)
    for (int Sy = 0; Sy <= Sy_max; Sy++)
      for (int Sz = 0; Sz <= Sz_max; Sz++) {
        const int lx = Sx - (Sx == 0 ? 0 : 1), ly = Sy - (Sy == 0 ? 0 : 1), lz = Sz - (Sz == 0 ? 0 : 1);
        const double R2_min = dx * dx * lx * lx + dy * dy * ly * ly + dz * dz * lz * lz;
        const double R2_max = dx * dx * (Sx + 1) * (Sx + 1) + dy * dy * (Sy + 1) * (Sy + 1) + dz * dz * (Sz + 1) * (Sz + 1);
        // 0 eta, 1 s, 2-4 v, 5 Txx, 6 Txy, 7 Txz, 8 Tyy, 9 Tyz, 10 Tzz
        double c[11] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (R * R > R2_max) {
          c[0] = dV;
        } else if (R * R > R2_min) {
          for (int ax : I) {
            int ix = ax;
            double Tx = dx * (Sx + ix);
            const int px = (Tx < 0 ? -1 : 1);
            if (Tx < 0) { ix = 1; Tx = dx; }
            for (int ay : I) {
              int iy = ay;
              double Ty = dy * (Sy + iy);
              const int py = (Ty < 0 ? -1 : 1);
              if (Ty < 0) { iy = 1; Ty = dy; }
              for (int az : I) {
                int iz = az;
                double Tz = dz * (Sz + iz);
                const int pz = (Tz < 0 ? -1 : 1);
                if (Tz < 0) { iz = 1; Tz = dz; }
                const int vx[2] = {0, ix}, vy[2] = {0, iy}, vz[2] = {0, iz};
                for (int jx = 0; jx < 2; jx++)
                  for (int jy = 0; jy < 2; jy++)
                    for (int jz = 0; jz < 2; jz++) {
                      const double Vx = Tx - dx * vx[jx], Vy = Ty - dy * vy[jy], Vz = Tz - dz * vz[jz];
                      int sgn = 1;
                      if (jx == 1) sgn *= -1;
                      if (jy == 1) sgn *= -1;
                      if (jz == 1) sgn *= -1;
                      if ((R * R - (Vx * Vx + Vy * Vy + Vz * Vz)) < std::nextafter(0.0, 1.0)) continue;
                      // upper limits as the reference writes them for each weight
                      const double hyz = std::sqrt(R * R - Vy * Vy - Vz * Vz);
                      const double hxz = std::sqrt(R * R - Vx * Vx - Vz * Vz);
                      const double hyx = std::sqrt(R * R - Vy * Vy - Vx * Vx);
                      const double hzx = std::sqrt(R * R - Vz * Vz - Vx * Vx);
                      c[0] += (1 / dV) * sgn * (ref_G(K_ETA, R, Vx, Vy, Vz, Tx, Ty, Tz) - ref_G(K_ETA, R, hyz, Vy, Vz, Tx, Ty, Tz));
                      c[1] += (1 / dV) * sgn * (ref_G(K_S, R, Vx, Vy, Vz, Tx, Ty, Tz) - ref_G(K_S, R, hyz, Vy, Vz, Tx, Ty, Tz));
                      c[2] += (1 / dV) * px * (1.0 / R) * sgn *
                              (ref_G(K_VX, R, Vx, Vy, Vz, Tx, Ty, Tz) - ref_G(K_VX, R, hyz, Vy, Vz, Tx, Ty, Tz));
                      c[3] += (1 / dV) * py * (1.0 / R) * sgn *
                              (ref_G(K_VX, R, Vy, Vx, Vz, Ty, Tx, Tz) - ref_G(K_VX, R, hxz, Vx, Vz, Ty, Tx, Tz));
                      c[4] += (1 / dV) * pz * (1.0 / R) * sgn *
                              (ref_G(K_VX, R, Vz, Vy, Vx, Tz, Ty, Tx) - ref_G(K_VX, R, hyx, Vy, Vx, Tz, Ty, Tx));
                      c[5] += (1 / dV) * (1.0 / (R * R)) * sgn *
                              (ref_G(K_TXX, R, Vx, Vy, Vz, Tx, Ty, Tz) - ref_G(K_TXX, R, hyz, Vy, Vz, Tx, Ty, Tz));
                      c[8] += (1 / dV) * (1.0 / (R * R)) * sgn *
                              (ref_G(K_TXX, R, Vy, Vx, Vz, Ty, Tx, Tz) - ref_G(K_TXX, R, hxz, Vx, Vz, Ty, Tx, Tz));
                      c[10] += (1 / dV) * (1.0 / (R * R)) * sgn *
                               (ref_G(K_TXX, R, Vz, Vy, Vx, Tz, Ty, Tx) - ref_G(K_TXX, R, hyx, Vy, Vx, Tz, Ty, Tx));
                      c[6] += (1 / dV) * px * py * (1.0 / (R * R)) * sgn *
                              (ref_G(K_TXY, R, Vx, Vy, Vz, Tx, Ty, Tz) - ref_G(K_TXY, R, hyz, Vy, Vz, Tx, Ty, Tz));
                      c[7] += (1 / dV) * px * pz * (1.0 / (R * R)) * sgn *
                              (ref_G(K_TXY, R, Vx, Vz, Vy, Tx, Tz, Ty) - ref_G(K_TXY, R, hyz, Vz, Vy, Tx, Tz, Ty));
                      c[9] += (1 / dV) * py * pz * (1.0 / (R * R)) * sgn *
                              (ref_G(K_TXY, R, Vy, Vz, Vx, Ty, Tz, Tx) - ref_G(K_TXY, R, hzx, Vz, Vx, Ty, Tz, Tx));
                    }
              }
            }
          }
        }
        // octant images with the parity of each component (src/FMT_Species.cpp:471-493), periodic wrap as
        // Lattice::putIntoBox (include/Lattice.h:154-164)
        for (int ox = 0; ox < (Sx == 0 ? 1 : 2); ox++)
          for (int oy = 0; oy < (Sy == 0 ? 1 : 2); oy++)
            for (int oz = 0; oz < (Sz == 0 ? 1 : 2); oz++) {
              int px = (1 - 2 * ox) * Sx, py = (1 - 2 * oy) * Sy, pz = (1 - 2 * oz) * Sz;
              while (px < 0) px += Nx;
              while (px >= Nx) px -= Nx;
              while (py < 0) py += Ny;
              while (py >= Ny) py -= Ny;
              while (pz < 0) pz += Nz;
              while (pz >= Nz) pz -= Nz;
              const long pos = pz + (long)Nz * (py + (long)Ny * px);
              const int s[3] = {1 - 2 * ox, 1 - 2 * oy, 1 - 2 * oz};
              w[0 * N + pos] += c[0];
              w[1 * N + pos] += c[1];
              w[2 * N + pos] += s[0] * c[2];
              w[3 * N + pos] += s[1] * c[3];
              w[4 * N + pos] += s[2] * c[4];
              w[5 * N + pos] += s[0] * s[0] * c[5];
              w[6 * N + pos] += s[0] * s[1] * c[6];
              w[7 * N + pos] += s[0] * s[2] * c[7];
              w[8 * N + pos] += s[1] * s[1] * c[8];
              w[9 * N + pos] += s[1] * s[2] * c[9];
              w[10 * N + pos] += s[2] * s[2] * c[10];
            }
      }
}

}  // namespace cdft
