// Species constraints of the full evaluation (Species::beginForceCalculation / endForceCalculation,
// src/Species.cpp:64-112): fixed mass (rescale rho, project dF), fixed background (dF = 0 on the boundary)
// and homogeneous boundary (boundary dF replaced by its mean).  Boundary points are those with ix, iy or
// iz = 0 (Lattice::is_boundary_point, include/Lattice.h:189-196); they are enumerated once each as the
// plane ix = 0, then iy = 0 with ix >= 1, then iz = 0 with ix, iy >= 1:
// Nb = Ny Nz + (Nx - 1) Nz + (Nx - 1)(Ny - 1) = NxNy + NxNz + NyNz - Nx - Ny - Nz + 1.
// All kernels take grid.y = batch member.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "rows2.hpp"

namespace cdft {

__host__ __device__ inline long boundary_count(int Nx, int Ny, int Nz) {
  return (long)Ny * Nz + (long)(Nx - 1) * Nz + (long)(Nx - 1) * (Ny - 1);
}

// p in [0, boundary_count) -> linear site index (z fastest)
__device__ inline long boundary_site(long p, int Nx, int Ny, int Nz) {
  const long a = (long)Ny * Nz;
  if (p < a) return p;  // ix = 0
  p -= a;
  const long b = (long)(Nx - 1) * Nz;
  if (p < b) return (1 + p / Nz) * a + p % Nz;  // iy = 0, ix >= 1
  p -= b;
  return ((1 + p / (Ny - 1)) * Ny + 1 + p % (Ny - 1)) * (long)Nz;  // iz = 0, ix, iy >= 1
}

enum { CONS_SUM = 0, CONS_DOT = 1, CONS_BSUM = 2, CONS_BABS = 3 };

// Compensated partial sums per block: sum a (CONS_SUM), sum a b (CONS_DOT), sum of a over the boundary
// (CONS_BSUM), sum |a| over the boundary (CONS_BABS; n = boundary count for the last two).
static __global__ __launch_bounds__(256) void k_cons_sum(const double* __restrict__ a, const double* __restrict__ b,
                                                         long n, long stride, int mode, int Nx, int Ny, int Nz,
                                                         double2* __restrict__ partials, long sp) {
  a += blockIdx.y * stride;
  if (b) b += blockIdx.y * stride;
  partials += blockIdx.y * sp;
  CSum acc{0.0, 0.0};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    if (mode == CONS_SUM) acc.add(a[i]);
    else if (mode == CONS_DOT) acc.add(a[i] * b[i]);
    else {
      const double v = a[boundary_site(i, Nx, Ny, Nz)];
      acc.add(mode == CONS_BABS ? fabs(v) : v);
    }
  }
  CSum v[1] = {acc};
  block_reduce_store_n<1>(v, partials, blockIdx.x);
}

// rho *= mass / (sum rho dV): Density::scale_to(fixedMass_ / getNumberAtoms()), src/Species.cpp:66-69
// (mass[] per member)
static __global__ __launch_bounds__(256) void k_cons_scale(double* __restrict__ rho, long n, long stride,
                                                           const double* __restrict__ res, long so,
                                                           const double* __restrict__ mass, double dV) {
  rho += blockIdx.y * stride;
  const double f = mass[blockIdx.y] / (res[blockIdx.y * so] * dV);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) rho[i] *= f;
}

// boundary values <- 0 (res == nullptr) or res * scale (src/Species.cpp:78-98)
static __global__ __launch_bounds__(256) void k_cons_bset(double* __restrict__ a, long nbnd, long stride, int Nx,
                                                          int Ny, int Nz, const double* __restrict__ res, long so,
                                                          double scale) {
  a += blockIdx.y * stride;
  const double v = res ? res[blockIdx.y * so] * scale : 0.0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < nbnd; p += (long)gridDim.x * blockDim.x)
    a[boundary_site(p, Nx, Ny, Nz)] = v;
}

// dF -= mu dV with mu = (sum dF rho) / mass (src/Species.cpp:99-110; mass[] per member)
static __global__ __launch_bounds__(256) void k_cons_shift(double* __restrict__ dF, long n, long stride,
                                                           const double* __restrict__ res, long so,
                                                           const double* __restrict__ mass, double dV) {
  dF += blockIdx.y * stride;
  const double mu = res[blockIdx.y * so] / mass[blockIdx.y];
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dF[i] -= mu * dV;
}

}  // namespace cdft
