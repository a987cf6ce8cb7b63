// Device generation of the cell-integrated FMT weights and of the LJ
// mean-field kernel (one-time setup per system).
//
// FMT weights: the exact integrals of the sphere (volume, surface, vector and
// tensor densities) against the trilinear hat of each lattice cell, as in
// FMT_Species::generateWeights (src/FMT_Species.cpp:310-496) with the G
// functions of :112-308.  The integrals are evaluated in a well-conditioned
// form: every asin(p/q) of the reference whose complementary leg is known
// exactly is written atan2(p, leg).  With E = R^2 - X^2 - V^2 the three angles
// are atan2(X, sqrt E), atan2(V, sqrt E) and atan2(X V, R sqrt E), and at the
// upper limit X = sqrt(R^2 - Va^2 - Vb^2) one has E_a = Vb^2, E_b = Va^2
// exactly.  The reference's asin form loses up to sqrt(eps) where an argument
// is exactly +-1 (a corner coordinate equal to zero); this form keeps the sum
// rules (sum w_eta = pi d^3/6, sum w_s = pi d^2, sum w_Tii = pi d^2/3) to
// ~1e-12 relative in double.
//
// One thread evaluates one cell (Sx,Sy,Sz) of the positive octant; a second
// kernel gathers the signed octant images onto the periodic grid.
#pragma once
#include "potential.hpp"
#include <hip/hip_runtime.h>

namespace cdft {

constexpr double kPiW = 3.14159265358979323846;

struct WMoments {
  double Iy[5], Iz[5], Jy[5], Jz[5];
};

__device__ __forceinline__ void wc_IJ(double X, double V, double R, double E, double* I, double* J) {
  if (E < 0) E = 0;
  const double sE = sqrt(E);
  const double aX = atan2(X, sE);
  const double aV = atan2(V, sE);
  const double aVX = atan2(X * V, R * sE);
  const double A2 = X * X + E;
  const double X2 = X * X, R2 = R * R, V2 = V * V;
  I[0] = 0.5 * X * sE + 0.5 * A2 * aX;
  I[1] = -(1.0 / 3) * E * sE;
  I[2] = 0.125 * X * (2 * X2 - A2) * sE + 0.125 * A2 * A2 * aX;
  I[3] = -(1.0 / 15) * E * (3 * X2 + 2 * A2) * sE;
  I[4] = (1.0 / 48) * (8 * X2 * X2 - 2 * A2 * X2 - 3 * A2 * A2) * X * sE + 0.0625 * A2 * A2 * A2 * aX;
  J[0] = X * aV + V * aX - R * aVX;
  J[1] = -0.5 * V * sE + 0.5 * (X2 - R2) * aV;
  J[2] = -(1.0 / 6) * V * (V2 - 3 * R2) * aX + (1.0 / 3) * X2 * X * aV - (1.0 / 3) * R2 * R * aVX - (1.0 / 6) * V * X * sE;
  J[3] = 0.25 * (X2 * X2 - R2 * R2) * aV + (kPiW / 8) * R2 * R2 - (V / 12.0) * (5 * R2 + X2 - 2 * V2) * sE;
  J[4] = 0.025 * X * V * (3 * V2 - 2 * X2 - 7 * R2) * sE + 0.025 * V * (3 * V2 * V2 - 10 * V2 * R2 + 15 * R2 * R2) * aX +
         0.2 * X2 * X2 * X * aV - 0.2 * R2 * R2 * R * aVX;
}

// kind: 0 eta, 1 s, 2 v (moment order 1), 3 txx (order 2), 4 txy
__device__ double wc_G(int kind, double R, double X, double Vy, double Vz, double Tx, double Ty, double Tz,
                       double Ey, double Ez) {
  WMoments q;
  wc_IJ(X, Vy, R, Ey, q.Iy, q.Jy);
  wc_IJ(X, Vz, R, Ez, q.Iz, q.Jz);
  const double R2 = R * R, X2 = X * X;
  if (kind == 0) {  // G_eta, FMT_Species.cpp:148-193
    const double A = Ty * Tz * Vy * Vz + 0.25 * Vy * Vy * Vz * Vz + 0.125 * R2 * R2 -
                     (1.0 / 6) * (Ty * Vy * Vy * Vy + Tz * Vz * Vz * Vz) - 0.5 * Tz * Vz * Vy * Vy -
                     0.5 * Ty * Vy * Vz * Vz + 0.125 * Vy * Vy * Vy * Vy + 0.125 * Vz * Vz * Vz * Vz +
                     0.5 * R2 * (Ty * Vy + Tz * Vz - 0.5 * Vy * Vy - 0.5 * Vz * Vz + 0.5 * kPiW * Ty * Tz);
    const double B = 0.5 * Ty * Vy + 0.5 * Tz * Vz - 0.25 * Vy * Vy - 0.25 * Vz * Vz + (kPiW / 4) * Ty * Tz + 0.25 * R2;
    const double C = 0.5 * Ty * Vy + (1.0 / 3) * (R2 - Vy * Vy);
    const double D = 0.5 * Tz * Vz + (1.0 / 3) * (R2 - Vz * Vz);
    double g = Tx * A * X - 0.5 * A * X2 - (1.0 / 3) * B * Tx * X2 * X + 0.25 * X2 * X2 * B + 0.025 * Tx * X2 * X2 * X -
               (1.0 / 48) * X2 * X2 * X2;
    g -= 0.5 * Ty * Tz * (Tx * R2 * q.Jy[0] - R2 * q.Jy[1] - Tx * q.Jy[2] + q.Jy[3]);
    g -= 0.5 * Ty * Tz * (Tx * R2 * q.Jz[0] - R2 * q.Jz[1] - Tx * q.Jz[2] + q.Jz[3]);
    g -= (Tx * Tz * C * q.Iy[0] - Tz * C * q.Iy[1] - (1.0 / 3) * Tx * Tz * q.Iy[2] + (1.0 / 3) * Tz * q.Iy[3]);
    g -= (Tx * Ty * D * q.Iz[0] - Ty * D * q.Iz[1] - (1.0 / 3) * Tx * Ty * q.Iz[2] + (1.0 / 3) * Ty * q.Iz[3]);
    return g;
  }
  if (kind == 4) {  // G_txy, FMT_Species.cpp:280-308
    const double A = (1.0 / 6) * (3 * Ty * R2 + 2 * Vy * Vy * Vy - 3 * Ty * Vy * Vy - 3 * Ty * Vz * Vz + 6 * Ty * Tz * Vz +
                                  1.5 * kPiW * R2 * Tz);
    const double B = 0.25 * kPiW * Tz + 0.5 * Ty;
    const double C = -(1.0 / 6) * (2 * R2 - 2 * Vz * Vz + 3 * Tz * Vz);
    double g = -0.5 * R * A * Tx * X2 + (1.0 / 3) * R * A * X2 * X + 0.25 * R * B * Tx * X2 * X2 - 0.2 * R * B * X2 * X2 * X;
    g += 0.5 * R * Tz * (Tx * R2 * q.Jy[1] - R2 * q.Jy[2] - Tx * q.Jy[3] + q.Jy[4]);
    g += 0.5 * R * Tz * (Tx * R2 * q.Jz[1] - R2 * q.Jz[2] - Tx * q.Jz[3] + q.Jz[4]);
    g += 0.5 * R * Tz * (2 * Ty - Vy) * (Tx * q.Iy[1] - q.Iy[2]);
    g += R * (-Tx * C * q.Iz[1] + C * q.Iz[2] - (1.0 / 3) * Tx * q.Iz[3] + (1.0 / 3) * q.Iz[4]);
    return g;
  }
  // G_s / G_vx / G_txx share one structure with moment order o (:195-278)
  const int o = kind - 1;
  const double A = Ty * Vy + Tz * Vz - 0.5 * Vy * Vy - 0.5 * Vz * Vz + 0.5 * R2 + (kPiW / 2) * Ty * Tz;
  double g;
  if (o == 0) g = 0.125 * R * X2 * X2 - (1.0 / 6) * R * Tx * X2 * X - 0.5 * R * A * X2 + R * A * Tx * X;
  else if (o == 1) g = 0.1 * R * X2 * X2 * X - 0.125 * R * Tx * X2 * X2 - (1.0 / 3) * R * A * X2 * X + 0.5 * R * A * Tx * X2;
  else g = (1.0 / 12) * R * X2 * X2 * X2 - 0.1 * R * Tx * X2 * X2 * X - 0.25 * R * A * X2 * X2 + (1.0 / 3) * R * A * Tx * X2 * X;
  g -= R * Tz * (Tx * q.Iy[o] - q.Iy[o + 1]);
  g -= R * Ty * (Tx * q.Iz[o] - q.Iz[o + 1]);
  g -= R * Ty * Tz * (Tx * q.Jy[o] - q.Jy[o + 1]);
  g -= R * Ty * Tz * (Tx * q.Jz[o] - q.Jz[o + 1]);
  return g;
}

// G(lower corner) - G(surface), roles already permuted by the caller.
__device__ __forceinline__ double wc_Gdiff(int kind, double R, double Xc, double Va, double Vb, double Tx, double Ta,
                                           double Tb) {
  const double R2 = R * R;
  const double h = sqrt(R2 - Va * Va - Vb * Vb);
  const double lo = wc_G(kind, R, Xc, Va, Vb, Tx, Ta, Tb, R2 - Xc * Xc - Va * Va, R2 - Xc * Xc - Vb * Vb);
  const double hi = wc_G(kind, R, h, Va, Vb, Tx, Ta, Tb, Vb * Vb, Va * Va);
  return lo - hi;
}

// One thread per cell (Sx,Sy,Sz) in [0,Smax]^3: writes the 11 cell values
// (eta, s, vx, vy, vz, Txx, Txy, Txz, Tyy, Tyz, Tzz) for the positive octant.
static __global__ void k_fmt_cells(int Sxm, int Sym, int Szm, double dx, double dy, double dz, double R, double* __restrict__ cells) {
  const int ncell = (Sxm + 1) * (Sym + 1) * (Szm + 1);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int Sz = c % (Szm + 1);
  const int Sy = (c / (Szm + 1)) % (Sym + 1);
  const int Sx = c / ((Szm + 1) * (Sym + 1));
  const double dV = dx * dy * dz;
  auto lo = [](int S) { return (double)(S == 0 ? 0 : S - 1); };
  const double R2min = dx * dx * lo(Sx) * lo(Sx) + dy * dy * lo(Sy) * lo(Sy) + dz * dz * lo(Sz) * lo(Sz);
  const double R2max = dx * dx * (Sx + 1.0) * (Sx + 1.0) + dy * dy * (Sy + 1.0) * (Sy + 1.0) + dz * dz * (Sz + 1.0) * (Sz + 1.0);
  double w[11];
  for (int a = 0; a < 11; ++a) w[a] = 0.0;
  if (R * R > R2max) {
    w[0] = dV;
  } else if (R * R > R2min) {
    for (int ax = -1; ax <= 1; ax += 2) {
      int ix = ax;
      double Tx = dx * (Sx + ix);
      const int px = Tx < 0 ? -1 : 1;
      if (Tx < 0) { ix = 1; Tx = dx; }
      for (int ay = -1; ay <= 1; ay += 2) {
        int iy = ay;
        double Ty = dy * (Sy + iy);
        const int py = Ty < 0 ? -1 : 1;
        if (Ty < 0) { iy = 1; Ty = dy; }
        for (int az = -1; az <= 1; az += 2) {
          int iz = az;
          double Tz = dz * (Sz + iz);
          const int pz = Tz < 0 ? -1 : 1;
          if (Tz < 0) { iz = 1; Tz = dz; }
          for (int j = 0; j < 8; ++j) {
            const int jx = j >> 2, jy = (j >> 1) & 1, jz = j & 1;
            const double Vx = Tx - dx * (jx ? ix : 0), Vy = Ty - dy * (jy ? iy : 0), Vz = Tz - dz * (jz ? iz : 0);
            if ((R * R - (Vx * Vx + Vy * Vy + Vz * Vz)) <= 0.0) continue;  // corner outside the sphere
            const double cs = (((jx + jy + jz) & 1) ? -1.0 : 1.0) / dV;
            const double cr = cs / (R * R);
            w[0] += cs * wc_Gdiff(0, R, Vx, Vy, Vz, Tx, Ty, Tz);
            w[1] += cs * wc_Gdiff(1, R, Vx, Vy, Vz, Tx, Ty, Tz);
            w[2] += cs * px / R * wc_Gdiff(2, R, Vx, Vy, Vz, Tx, Ty, Tz);
            w[3] += cs * py / R * wc_Gdiff(2, R, Vy, Vx, Vz, Ty, Tx, Tz);
            w[4] += cs * pz / R * wc_Gdiff(2, R, Vz, Vy, Vx, Tz, Ty, Tx);
            w[5] += cr * wc_Gdiff(3, R, Vx, Vy, Vz, Tx, Ty, Tz);
            w[8] += cr * wc_Gdiff(3, R, Vy, Vx, Vz, Ty, Tx, Tz);
            w[10] += cr * wc_Gdiff(3, R, Vz, Vy, Vx, Tz, Ty, Tx);
            w[6] += cr * (px * py) * wc_Gdiff(4, R, Vx, Vy, Vz, Tx, Ty, Tz);
            w[7] += cr * (px * pz) * wc_Gdiff(4, R, Vx, Vz, Vy, Tx, Tz, Ty);
            w[9] += cr * (py * pz) * wc_Gdiff(4, R, Vy, Vz, Vx, Ty, Tz, Tx);
          }
        }
      }
    }
  }
  for (int a = 0; a < 11; ++a) cells[(long)c * 11 + a] = w[a];
}

// Gather the signed octant images of the cells onto the periodic grid
// (octant scatter of FMT_Species.cpp:471-493, written as a gather so every
// grid point is owned by one thread).  out: 11 arrays of N doubles.
static __global__ void k_fmt_gather(int Nx, int Ny, int Nz, int Sxm, int Sym, int Szm, const double* __restrict__ cells,
                             double* __restrict__ out) {
  const long N = (long)Nx * Ny * Nz;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int z = p % Nz, y = (p / Nz) % Ny, x = p / ((long)Nz * Ny);
  // candidate (S, sign) per axis: S*sign == coordinate (mod N), 0 <= S <= Smax
  int cx[8], sx[8], nxc = 0, cy[8], sy[8], nyc = 0, cz[8], sz[8], nzc = 0;
  for (int S = 0; S <= Sxm && nxc < 8; ++S) {
    if (S % Nx == x && nxc < 8) { cx[nxc] = S; sx[nxc++] = 1; }
    if (S > 0 && ((Nx - S % Nx) % Nx) == x && nxc < 8) { cx[nxc] = S; sx[nxc++] = -1; }
  }
  for (int S = 0; S <= Sym && nyc < 8; ++S) {
    if (S % Ny == y && nyc < 8) { cy[nyc] = S; sy[nyc++] = 1; }
    if (S > 0 && ((Ny - S % Ny) % Ny) == y && nyc < 8) { cy[nyc] = S; sy[nyc++] = -1; }
  }
  for (int S = 0; S <= Szm && nzc < 8; ++S) {
    if (S % Nz == z && nzc < 8) { cz[nzc] = S; sz[nzc++] = 1; }
    if (S > 0 && ((Nz - S % Nz) % Nz) == z && nzc < 8) { cz[nzc] = S; sz[nzc++] = -1; }
  }
  double acc[11];
  for (int a = 0; a < 11; ++a) acc[a] = 0.0;
  for (int i = 0; i < nxc; ++i)
    for (int j = 0; j < nyc; ++j)
      for (int k = 0; k < nzc; ++k) {
        const long c = ((long)cx[i] * (Sym + 1) + cy[j]) * (Szm + 1) + cz[k];
        const double* w = cells + c * 11;
        const double g[3] = {(double)sx[i], (double)sy[j], (double)sz[k]};
        acc[0] += w[0];
        acc[1] += w[1];
        acc[2] += g[0] * w[2];
        acc[3] += g[1] * w[3];
        acc[4] += g[2] * w[4];
        acc[5] += w[5];
        acc[6] += g[0] * g[1] * w[6];
        acc[7] += g[0] * g[2] * w[7];
        acc[8] += w[8];
        acc[9] += g[1] * g[2] * w[9];
        acc[10] += w[10];
      }
  for (int a = 0; a < 11; ++a) out[a * N + p] = acc[a];
}

// LJ with WCA split: Potential1::Watt2 (include/Potential1.h:56-65) for LJ
// (:136-189), bhFlag = false, r_att_min = 0.
// Lattice mean-field kernel (Interaction_Base::generateWeightsXYZSym, src/Interaction.cpp:178-307, and
// initialize :32-52): w(S) = dx^2 dy^2 dz^2 generateWeight(S) with the offsets sorted |Sx| >= |Sy| >= |Sz| as the
// reference tabulates them, gathered onto the periodic grid over |offset| <= lim per axis, divided by dV and kT.
static __global__ void k_mf_kernel(int Nx, int Ny, int Nz, double dx, double dy, double dz, int lx, int ly, int lz, Pot v,
                                   MFRule rule, double kT, double* __restrict__ out) {
  const long N = (long)Nx * Ny * Nz;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int z = p % Nz, y = (p / Nz) % Ny, x = p / ((long)Nz * Ny);
  const double gf = dx * dx * dy * dy * dz * dz;
  const double dV = dx * dy * dz;
  double acc = 0.0;
  // smallest offset congruent to the coordinate, then step by the period
  int ix0 = x - ((x + lx) / Nx) * Nx;
  int iy0 = y - ((y + ly) / Ny) * Ny;
  int iz0 = z - ((z + lz) / Nz) * Nz;
  for (int ix = ix0; ix <= lx; ix += Nx)
    for (int iy = iy0; iy <= ly; iy += Ny)
      for (int iz = iz0; iz <= lz; iz += Nz) {
        int a = ix < 0 ? -ix : ix, b = iy < 0 ? -iy : iy, c = iz < 0 ? -iz : iz, t;
        if (b > a) { t = a; a = b; b = t; }
        if (c > a) { t = a; a = c; c = t; }
        if (c > b) { t = b; b = c; c = t; }
        acc += gf * mf_cell_weight(rule, v, a, b, c, dx, dy, dz) / dV;
      }
  out[p] = acc / kT;
}

// Parity of the weight arrays in x (bit a set = odd): 2 v_x, 6 T_xy, 7 T_xz (include/Species.h:351-360 order); eta,
// s, v_y, v_z, the other tensor entries and the mean-field kernel (stored after the m FMT arrays, index 5 or 11)
// are even in x.  kOddY is the same for y (used by the check only).
constexpr unsigned kOddX = (1u << 2) | (1u << 6) | (1u << 7);
constexpr unsigned kOddY = (1u << 3) | (1u << 6) | (1u << 9);

// Symmetry check of real spectra W (narr arrays of Nx x Ny x Zp): per block and array, max |W(k) - s_x W(-kx)|,
// max |W(k) - s_y W(-ky)| and max |W(k)| over kz < Zh.  grid = (nblk, narr).
template <typename T>
__global__ __launch_bounds__(256) void k_wsym(const T* __restrict__ W, long nkp, int Nx, int Ny, int Zh, int Zp,
                                              unsigned pxm, unsigned pym, double* __restrict__ out) {
  const int a = blockIdx.y;
  W += (long)a * nkp;
  const long n = (long)Nx * Ny * Zh;
  double dev = 0.0, mx = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int kz = (int)(i % Zh), ky = (int)((i / Zh) % Ny), kx = (int)(i / ((long)Zh * Ny));
    const double v = (double)W[((long)kx * Ny + ky) * Zp + kz];
    const double vx = (double)W[((long)((Nx - kx) % Nx) * Ny + ky) * Zp + kz];
    const double vy = (double)W[((long)kx * Ny + (Ny - ky) % Ny) * Zp + kz];
    dev = fmax(dev, fabs(v - (((pxm >> a) & 1u) ? -vx : vx)));
    dev = fmax(dev, fabs(v - (((pym >> a) & 1u) ? -vy : vy)));
    mx = fmax(mx, fabs(v));
  }
  __shared__ double sd[256], sm[256];
  sd[threadIdx.x] = dev;
  sm[threadIdx.x] = mx;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + off]);
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * ((long)a * gridDim.x + blockIdx.x)] = sd[0];
    out[2 * ((long)a * gridDim.x + blockIdx.x) + 1] = sm[0];
  }
}

// Wq[a][(kx Ny + ky) Zp + kz] = W[a][(kx Ny + ky) Zp + kz] for kx <= Nx/2 (the first rows of each array).
template <typename T>
__global__ __launch_bounds__(256) void k_wfold(const T* __restrict__ W, T* __restrict__ Wq, long nkp, long nq, int narr) {
  const long n = nq * narr;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long a = i / nq, r = i - a * nq;
    Wq[i] = W[a * nkp + r];
  }
}

// Points-file mean-field kernel (Interaction::generateWeights, src/Interaction.cpp:672-860): every direction u of
// the points list (unit vector, weight), under the 6 axis permutations and 8 sign reflections, and every radial
// Gauss-Legendre node r on [0, Rc] puts w = weight / 48 * wr r^2 Watt(r) on the 8 corners of the lattice cell
// holding r u, with trilinear shares (periodic wrap).  One thread per (direction, permutation, reflection,
// radial node); the corners accumulate with atomic adds, so the last bits of a cell depend on the order.
// The kernel is returned before the division by kT, as the reference's weight cache stores it.
static __global__ __launch_bounds__(256) void k_mf_points(const double* __restrict__ pts, long npts,
                                                          const double* __restrict__ rr, const double* __restrict__ wr,
                                                          int nr, Pot v, int Nx, int Ny, int Nz, double dx, double dy,
                                                          double dz, double* __restrict__ out) {
  const long total = npts * 48 * nr;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int k = (int)(t % nr);
    const int is = (int)((t / nr) % 8);
    const int iperm = (int)((t / ((long)nr * 8)) % 6);
    const long pos = t / ((long)nr * 48);
    // permutation (ii, jj, kk) of the direction's components, src/Interaction.cpp:780-791
    const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    const double* u = pts + 4 * pos;
    const double r = rr[k];
    const double watt = v.watt2(r * r);
    // reflections, :802-811: x flips for is in {1,4,5,7}, y for {2,4,6,7}, z for {3,5,6,7}, as bit masks over is
    // (0xB2, 0xD4, 0xE8) and exact multiplications by -1; the equivalent chain of ifs was compiled for gfx950 with
    // the is == 3 case of the z flip dropped
    const double sx = ((0xB2 >> is) & 1) ? -1.0 : 1.0, sy = ((0xD4 >> is) & 1) ? -1.0 : 1.0,
                 sz = ((0xE8 >> is) & 1) ? -1.0 : 1.0;
    const double x = sx * (r * u[perm[iperm][0]]), y = sy * (r * u[perm[iperm][1]]), z = sz * (r * u[perm[iperm][2]]);
    const double ww = u[3] / 48.0;
    int ix0 = int(x / dx), iy0 = int(y / dy), iz0 = int(z / dz);
    if (x < 0) ix0 -= 1;
    if (y < 0) iy0 -= 1;
    if (z < 0) iz0 -= 1;
    const double ddx = (x / dx) - ix0, ddy = (y / dy) - iy0, ddz = (z / dz) - iz0;
    for (int io = 0; io < 2; ++io)
      for (int jo = 0; jo < 2; ++jo)
        for (int ko = 0; ko < 2; ++ko) {
          int nx = (ix0 + io) % Nx, ny = (iy0 + jo) % Ny, nz = (iz0 + ko) % Nz;
          if (nx < 0) nx += Nx;
          if (ny < 0) ny += Ny;
          if (nz < 0) nz += Nz;
          const double cc = (io == 0 ? 1 - ddx : ddx) * (jo == 0 ? 1 - ddy : ddy) * (ko == 0 ? 1 - ddz : ddz);
          atomicAdd(out + (nz + (long)Nz * (ny + (long)Ny * nx)), cc * ww * wr[k] * r * r * watt);
        }
  }
}

}  // namespace cdft
