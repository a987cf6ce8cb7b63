// Device generation of the cell-integrated FMT weights and of the LJ
// mean-field kernel (one-time setup per system).
//
// FMT weights: the exact integrals of the sphere (volume, surface, vector and
// tensor densities) against the trilinear hat of each lattice cell, as in
// FMT_Species::generateWeights (src/FMT_Species.cpp:310-496) with the G
// functions of :112-308, evaluated in the well-conditioned double-double form
// of wcells.hpp (cell weights correct to the last bit of a double).  One
// thread evaluates one weight of one cell (Sx,Sy,Sz) of the positive octant; a
// second kernel gathers the signed octant images onto the periodic grid.
#pragma once
#include "potential.hpp"
#include "wcells.hpp"
#include <hip/hip_runtime.h>

namespace cdft {

// Thread t: weight a = t / ncell of cell c = t % ncell, written to cells[c * 11 + a] for (Sx,Sy,Sz) in [0,Smax]^3
// with c = (Sx (Sym+1) + Sy)(Szm+1) + Sz.
static __global__ __launch_bounds__(64) void k_fmt_cells(int Sxm, int Sym, int Szm, double dx, double dy, double dz,
                                                         double R, double* __restrict__ cells) {
  const long ncell = (long)(Sxm + 1) * (Sym + 1) * (Szm + 1);
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 11 * ncell) return;
  const int a = (int)(t / ncell);
  const long c = t - a * ncell;
  const int Sz = (int)(c % (Szm + 1));
  const int Sy = (int)((c / (Szm + 1)) % (Sym + 1));
  const int Sx = (int)(c / ((long)(Szm + 1) * (Sym + 1)));
  cells[c * 11 + a] = wc_cell<dd>(a, Sx, Sy, Sz, dx, dy, dz, R);
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
