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
// the six independent tensor entries xx, xy, xz, yy, yz, zz), 10 without the tensor.  Each species' weighted
// densities come from its own forward convolution (x and y transformed back, z spectral); because the measure
// sums and the collapse are linear with real coefficients they are formed element by element in that space.
#pragma once
#include "rows2.hpp"

namespace cdft {

constexpr int kMixMeas = 16;

// Mx (nm arrays, astride apart) = (first) or += the measure contributions of one species' weighted-density arrays
// Bs (11 or 5 arrays, astride apart; FMT_Species index order, include/Species.h:351-360), for elements [0, n).
__global__ __launch_bounds__(256) void k_mix_accum(const cplx<double>* __restrict__ Bs, cplx<double>* __restrict__ Mx,
                                                   long astride, long n, double d, int tensor, int first) {
  const long step = (long)gridDim.x * blockDim.x;
  const double id = 1.0 / d, id2 = id * id;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
    auto put = [&](int a, double sc, cplx<double> v) {
      cplx<double> o = first ? cplx<double>{0.0, 0.0} : Mx[a * astride + i];
      o.x += sc * v.x;
      o.y += sc * v.y;
      Mx[a * astride + i] = o;
    };
    put(0, 1.0, Bs[i]);  // eta
    const cplx<double> s = Bs[astride + i];
    put(1, id2, s);  // s0 = A1 / d^2
    put(2, id, s);   // s1 = A1 / d
    put(3, 1.0, s);  // s2 = A1
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const cplx<double> v = Bs[(2 + j) * astride + i];
      put(4 + j, id, v);   // v1 = A_v / d
      put(7 + j, 1.0, v);  // v2 = A_v
    }
    if (tensor)
#pragma unroll
      for (int k = 0; k < 6; ++k) put(10 + k, 1.0, Bs[(5 + k) * astride + i]);
  }
}

// Bs (the species' 11 or 5 arrays) = the measure derivatives dM (nm arrays) collapsed with the species' diameter d
// (include/Species.h:310-329): eta; s0/d^2 + s1/d + s2; -(v2 + v1/d) (the parity sign of the odd vector weights);
// the tensor entries as stored by the pointwise kernel (off-diagonals already counted twice).
__global__ __launch_bounds__(256) void k_mix_collapse(const cplx<double>* __restrict__ dM, cplx<double>* __restrict__ Bs,
                                                      long astride, long n, double d, int tensor) {
  const long step = (long)gridDim.x * blockDim.x;
  const double id = 1.0 / d, id2 = id * id;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
    Bs[i] = dM[i];
    const cplx<double> a = dM[astride + i], b = dM[2 * astride + i], c = dM[3 * astride + i];
    Bs[astride + i] = {a.x * id2 + b.x * id + c.x, a.y * id2 + b.y * id + c.y};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const cplx<double> v1 = dM[(4 + j) * astride + i], v2 = dM[(7 + j) * astride + i];
      Bs[(2 + j) * astride + i] = {-(v2.x + v1.x * id), -(v2.y + v1.y * id)};
    }
    if (tensor)
#pragma unroll
      for (int k = 0; k < 6; ++k) Bs[(5 + k) * astride + i] = dM[(10 + k) * astride + i];
  }
}

// Pointwise z pass of the mixture (FP64): inverse z of the nm measure rows, Phi and its derivatives with respect to
// the measures per site (FMT::calculate_Phi / calculate_dPhi_wrt_fundamental_measures, src/FMT.cpp:114-176), the
// derivative rows transformed back in place; tensor off-diagonal derivatives are stored twice (the T_jk and T_kj
// terms, include/Species.h:327).  Sequence a * RB + r is measure a of row r; grid = ceil(nrows / RB),
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
    const double v1[3] = {n[4], n[5], n[6]}, v2[3] = {n[7], n[8], n[9]};
    PhiDeriv<double> d;
    const double phi = phi_meas<MODEL>(n[0], n[1], n[2], n[3], v1, v2, n + 10, pp, d);
    double out[kMixMeas] = {d.eta, d.s0, d.s1, d.s2, d.v1[0], d.v1[1], d.v1[2], d.v2[0], d.v2[1], d.v2[2],
                            d.T[0], 2 * d.T[1], 2 * d.T[2], d.T[3], 2 * d.T[4], d.T[5]};
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
