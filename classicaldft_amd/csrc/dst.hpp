// DST-I (FFTW RODFT00, Y_k = 2 sum_j X_j sin(pi (j+1)(k+1)/(n+1))) along one axis of a compact
// (n0, n1, n2) array (z fastest), and the element kernels of the open-boundary DDFT
// (DDFT_IF_Open, mduran/src/DDFT_IF_Open.cpp:25-463).
//
// A line of length n becomes the odd sequence y of length L = 2(n+1): y_0 = y_{n+1} = 0, y_m = x_{m-1},
// y_{L-m} = -x_{m-1}; its forward DFT is Y_k = -2i sum_m x_{m-1} sin(pi m k/(n+1)), so the DST-I is
// -Im Y_{k+1}.  Each block holds nbl complete sequences in LDS (nbl * L <= 256 * kDstPts) and runs the
// mixed-radix Stockham FFT of fft.hpp on them.  Lines along y and x are read with consecutive threads on
// consecutive lines (coalesced), lines along z with consecutive threads on consecutive elements.
#pragma once
#include <hip/hip_runtime.h>

#include "fft.hpp"

namespace cdft {

constexpr int kDstPts = 16;

__device__ __forceinline__ long dst_addr(long line, int j, int n1, int n2, int axis) {
  if (axis == 2) return line * n2 + j;
  if (axis == 1) return (line / n2) * (long)n1 * n2 + (long)j * n2 + line % n2;
  return (long)j * n1 * n2 + line;  // line = iy n2 + iz
}

static __global__ __launch_bounds__(256) void k_dst1(double* __restrict__ a, int n0, int n1, int n2, int axis,
                                                     Radices rad, const cplx<double>* __restrict__ tw, int nbl) {
  extern __shared__ unsigned char smem_dst[];
  cplx<double>* buf = reinterpret_cast<cplx<double>*>(smem_dst);
  const int n = axis == 0 ? n0 : (axis == 1 ? n1 : n2);
  const int L = rad.L;
  const long nlines = (long)n0 * n1 * n2 / n;
  const long line0 = (long)blockIdx.x * nbl;
  const int nb = (int)(nlines - line0 < nbl ? nlines - line0 : nbl);
  for (int i = threadIdx.x; i < nbl * L; i += blockDim.x) buf[i] = {0.0, 0.0};
  __syncthreads();
  for (int idx = threadIdx.x; idx < nb * n; idx += blockDim.x) {
    const int b = axis == 2 ? idx / n : idx % nb, j = axis == 2 ? idx % n : idx / nb;
    const double v = a[dst_addr(line0 + b, j, n1, n2, axis)];
    buf[b * L + j + 1] = {v, 0.0};
    buf[b * L + L - 1 - j] = {-v, 0.0};
  }
  __syncthreads();
  lds_fft<double, kDstPts, true, false>(buf, rad, nbl, 1, L, tw, threadIdx.x, blockDim.x);
  for (int idx = threadIdx.x; idx < nb * n; idx += blockDim.x) {
    const int b = axis == 2 ? idx / n : idx % nb, j = axis == 2 ? idx % n : idx / nb;
    a[dst_addr(line0 + b, j, n1, n2, axis)] = -buf[b * L + j + 1].y;
  }
}

// pack_for_sin_transform (:341-351): interior sites (ix, iy, iz >= 1) of x minus val, compact
static __global__ __launch_bounds__(256) void k_open_pack(const double* __restrict__ x, double* __restrict__ o, int Nx,
                                                          int Ny, int Nz, double val) {
  const int n1 = Ny - 1, n2 = Nz - 1;
  const long ni = (long)(Nx - 1) * n1 * n2;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < ni; p += (long)gridDim.x * blockDim.x) {
    const long iz = p % n2, iy = (p / n2) % n1, ix = p / ((long)n1 * n2);
    o[p] = x[((ix + 1) * Ny + iy + 1) * Nz + iz + 1] - val;
  }
}

// fftDiffusion's spectral update (:296-313): x <- x fac + (fac-1)/Lam R0 + (fac-1-dt Lam)/(Lam^2 dt) (R1 - R0)
// lam / ex: the per-axis eigenvalues and exp(lam dt), x then y then z
static __global__ __launch_bounds__(256) void k_open_update(double* __restrict__ S, const double* __restrict__ R0,
                                                            const double* __restrict__ R1,
                                                            const double* __restrict__ lam,
                                                            const double* __restrict__ ex, int n0, int n1, int n2,
                                                            double dt) {
  const long ni = (long)n0 * n1 * n2;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < ni; p += (long)gridDim.x * blockDim.x) {
    const int iz = (int)(p % n2), iy = (int)((p / n2) % n1), ix = (int)(p / ((long)n1 * n2));
    const double Lam = lam[ix] + lam[n0 + iy] + lam[n0 + n1 + iz];
    const double fac = ex[ix] * ex[n0 + iy] * ex[n0 + n1 + iz];
    double x = S[p] * fac;
    x += ((fac - 1) / Lam) * R0[p];
    x += ((fac - 1 - dt * Lam) / (Lam * Lam * dt)) * (R1[p] - R0[p]);
    S[p] = x;
  }
}

}  // namespace cdft
