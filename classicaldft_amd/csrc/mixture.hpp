// Hard-sphere mixtures: the FMT free energy and forces of several species of different diameters on one lattice.
//
// Reference: FMT::calculateFreeEnergyAndDerivatives over all species (src/FMT.cpp:363-427).  The fundamental
// measures at a site sum the weighted densities of every species with that species' diameter d
// (FMT::getWeightedDensities, src/FMT.cpp:78-111: eta += A0, s0 += A1/d^2, s1 += A1/d, s2 += A1, v1 += A_v/d,
// v2 += A_v, T += A_T), Phi and its derivatives are taken once per site, and each species collapses them with its
// own diameter (FMT_Species::set_fundamental_measure_derivatives, include/Species.h:310-329) before its own
// back-convolution (FMT_Species::Accumulate_dPhi, include/Species.h:172-178).
//
// Unlike one species, the measures s0, s1, s2 and v1, v2 are no longer multiples of one array, so the mixture
// carries them separately: kMixMeas = 16 measure arrays for the tensor models (eta, s0, s1, s2, v1[3], v2[3] and
// the six independent tensor entries xx, xy, xz, yy, yz, zz), 10 without the tensor.  Because the measure sums and
// the collapse are linear with real coefficients, they are folded into per-species measure weight spectra: each
// species' x-convolution writes (or adds to) the measures directly, and each species' backward x-convolution reads
// the measure derivatives directly.
#pragma once
#include "rows2.hpp"

namespace cdft {

constexpr int kMixMeas = 16;

// The measures as linear maps of one species' weight spectra (FMT::getWeightedDensities, src/FMT.cpp:78-111, and the
// collapse of FMT_Species::set_fundamental_measure_derivatives, include/Species.h:310-329): measure b is
// fwd_coef(b) times the weighted density of array mix_src(b), and contributes bwd_coef(b) times its derivative to
// the force through that array's weight.  Both maps are real and per site, so they commute with the transforms: a
// species' measures are its density spectrum times the weight spectra scaled by fwd_coef (one x-convolution per
// species, summed over species), and its force is the sum over measures of the derivative spectra times the weight
// spectra scaled by bwd_coef (one backward x-convolution per species).
__host__ __device__ constexpr int mix_src(int b) {
  return b == 0 ? 0 : b < 4 ? 1 : b < 7 ? b - 2 : b < 10 ? b - 5 : b - 5;
}
inline double mix_fwd_coef(int b, double d) {
  const double id = 1.0 / d;
  switch (b) {
    case 1: return id * id;                // s0 = A1 / d^2
    case 2: case 4: case 5: case 6: return id;  // s1 = A1 / d, v1 = A_v / d
    default: return 1.0;                   // eta, s2, v2, T
  }
}
inline double mix_bwd_coef(int b, double d) {
  const double id = 1.0 / d;
  if (b >= 4 && b < 7) return -id;   // -(v1 / d): the parity sign of the odd vector weights
  if (b >= 7 && b < 10) return -1.0;  // -v2
  return mix_fwd_coef(b, d);
}
// Measures whose weight is purely imaginary (the vectors) and those odd in x (v1_x, v2_x, T_xy, T_xz).
constexpr unsigned kMixOdd = 0x3F0u;
constexpr unsigned kMixOddX = (1u << 4) | (1u << 7) | (1u << 11) | (1u << 12);

// dst[i] = c * (s[i] - xx[i] - yy[i]), i < n: the T_zz weight of a species stored without it (the trace identity
// w_Txx + w_Tyy + w_Tzz = w_s), for a mixture that carries the T_zz measure.
__global__ __launch_bounds__(256) void k_mix_tzz(const double* __restrict__ ws, const double* __restrict__ xx,
                                                 const double* __restrict__ yy, double* __restrict__ dst, long n,
                                                 double c) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = c * (ws[i] - xx[i] - yy[i]);
}

// dst[i] = c * src[i], i < n.
__global__ __launch_bounds__(256) void k_mix_scale(const double* __restrict__ src, double* __restrict__ dst, long n,
                                                   double c) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = c * src[i];
}

// Pointwise z pass of the mixture (FP64): inverse z of the nm measure rows, Phi and its derivatives with respect to
// the measures per site (FMT::calculate_Phi / calculate_dPhi_wrt_fundamental_measures, src/FMT.cpp:114-176), the
// derivative rows transformed back in place; tensor off-diagonal derivatives are stored twice (the T_jk and T_kj
// terms, include/Species.h:327).  nm = 15 (every species' weights satisfy the trace identity): the T_zz measure is
// s2 - T_xx - T_yy per site and its derivative is folded back onto those three.  Sequence a * RB + r is measure a of row r; grid = ceil(nrows / RB),
// block = nm RB P, dynamic LDS = nm RB SLOT elements.  F partial sums at partials[block]; flag set when
// eta > 1 somewhere (src/FMT.cpp:289-290).
template <int MODEL, int M>
__global__ __launch_bounds__(256) void k_z_pw_mix(cplx<double>* __restrict__ Mx, long arr_stride, int nm, long nrows,
                                                  int Zp, const cplx<double>* __restrict__ twP,
                                                  const cplx<double>* __restrict__ twN, int RB, PhiParams pp,
                                                  double2* __restrict__ partials, int* __restrict__ flag) {
  using RF = RowF<double, M>;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
  const int seq = threadIdx.x / RF::P, t = threadIdx.x % RF::P;
  const int a = seq / RB, r = seq % RB;
  cplx<double>* slot = lds + seq * RF::SLOT;
  const long row0 = (long)blockIdx.x * RB;
  const long rrow = row0 + r < nrows ? row0 + r : nrows - 1;
  const bool ok = row0 + r < nrows;
  cplx<double>* brow = Mx + a * arr_stride + rrow * Zp;
  {
    cplx<double> v1[RF::Q1][RF::R1], v2[RF::Q2][RF::R2];
    c2r_to_regs<double, M>(brow, twN, t, v1);
    reg_fft_seqp<double, M, RF::R1, RF::R2, false, 1>(v1, v2, t, slot, 1, twP);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RF::Q2; ++q)
#pragma unroll
      for (int k2 = 0; k2 < RF::R2; ++k2) slot[t + RF::P * q + RF::R1 * k2] = v2[q][k2];
  }
  __syncthreads();
  CSum F{0.0, 0.0};
  int bad = 0;
  for (int s = threadIdx.x; s < RB * 2 * M; s += blockDim.x) {
    const int rr = s / (2 * M), z = s % (2 * M);
    if (row0 + rr >= nrows) continue;
    double n[kMixMeas];
#pragma unroll
    for (int b = 0; b < kMixMeas; ++b)
      n[b] = b < nm ? reinterpret_cast<const double*>(lds + (b * RB + rr) * RF::SLOT)[z] : 0.0;
    if (nm == 15) n[15] = n[3] - n[10] - n[13];
    const double v1[3] = {n[4], n[5], n[6]}, v2[3] = {n[7], n[8], n[9]};
    PhiDeriv<double> d;
    const double phi = phi_meas<MODEL>(n[0], n[1], n[2], n[3], v1, v2, n + 10, pp, d);
    double out[kMixMeas] = {d.eta, d.s0, d.s1, d.s2, d.v1[0], d.v1[1], d.v1[2], d.v2[0], d.v2[1], d.v2[2],
                            d.T[0], 2 * d.T[1], 2 * d.T[2], d.T[3], 2 * d.T[4], d.T[5]};
    if (nm == 15) {
      out[3] += out[15];
      out[10] -= out[15];
      out[13] -= out[15];
    }
    if (n[0] > 0.5 && 1 - n[0] < 0.0) bad = 1;
    F.add(phi);
#pragma unroll
    for (int b = 0; b < kMixMeas; ++b)
      if (b < nm) reinterpret_cast<double*>(lds + (b * RB + rr) * RF::SLOT)[z] = out[b];
  }
  if (bad) atomicOr(flag, 1);
  __syncthreads();
  {
    cplx<double> v1[RF::Q1][RF::R1], v2[RF::Q2][RF::R2];
#pragma unroll
    for (int q = 0; q < RF::Q1; ++q)
#pragma unroll
      for (int n1 = 0; n1 < RF::R1; ++n1) v1[q][n1] = slot[RF::R2 * n1 + RF::P * q + t];
    __syncthreads();
    reg_fft_seqp<double, M, RF::R1, RF::R2, true, 1>(v1, v2, t, slot, 1, twP);
    __syncthreads();
    r2c_from_regs<double, M>(slot, t, twN, brow, ok, v2);
  }
  CSum v[1] = {F};
  block_reduce_store_n<1>(v, partials, blockIdx.x);
}

}  // namespace cdft
