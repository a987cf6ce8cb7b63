// z-pass kernels (real rows <-> half spectra along z) built on the register
// FFT of fft2.hpp.  Same data layout and results as k_rows_r2c, k_rows_c2r,
// k_z_pointwise and k_z_final in kernels.hpp.
//
// A real row of length Nz = 2M is handled as one complex sequence of length M
// (z_m = x_2m + i x_2m+1) by P = gcd(R1, R2) threads (lane layout: thread
// fastest, then sequence).  The c2r pre-processing reads X_j and X_{M-j}
// straight from HBM into the pass-1 registers; the r2c post-processing pairs
// Z_k with Z_{M-k} through the sequence's LDS slot.  Each sequence owns a slot
// of R1 (R2 + 1) elements: the padded exchange of the transform, or M values in
// natural order.  twP: the length-M pass-1 twiddles in pass order
// (upload_pass_twiddles), twN: W_Nz^j for j <= M.
#pragma once
#include "kernels2.hpp"
#include "phi.hpp"

namespace cdft {

// Slot stride: the padded exchange R1 (R2 + 1), rounded up so that consecutive slots start an odd multiple of
// u = P esz bytes apart modulo the LDS's 256 B bank window (64 banks of 4 B).  An LDS access serves the lanes of
// 256 B per cycle: the 256 / u sequences of those lanes, each touching P consecutive elements (or, in the padded
// pass-2 reads, P elements R2 + 1 apart), then fill the window without sharing a bank.
constexpr int slot_stride(int base, int esz, int P) {
  const int u = P * esz;
  if (u >= 256) return base;
  for (int p = 0; p < 256; ++p) {
    const int r = (base + p) * esz % 256;
    if (r % u == 0 && (r / u) % 2 == 1) return base + p;
  }
  return base;
}

template <typename T, int M>
struct RowF {
  using F = F2<M>;
  static constexpr int R1 = F::R1, R2 = F::R2, P = F::P, Q1 = F::Q1, Q2 = F::Q2;
  static constexpr int SLOT = slot_stride(R1 * (R2 + 1), int(sizeof(cplx<T>)), P);
};

// Deterministic block reduction of NQ compensated sums for blockDim.x <= 1024:
// a fixed xor butterfly inside each wavefront (lanes beyond the block's
// thread count contribute zero), then thread 0 merges the wavefront results
// in order and writes partials[blk*NQ+q].
template <int NQ>
__device__ void block_reduce_store_n(CSum* v, double2* __restrict__ partials, long blk) {
  __shared__ double ws[NQ][16][2];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nact = min(64, (int)blockDim.x - wid * 64);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const bool have = (lane ^ off) < nact;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double s2 = __shfl_xor(v[q].s, off);
      const double c2 = __shfl_xor(v[q].c, off);
      if (have) v[q].merge(s2, c2);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      ws[q][wid][0] = v[q].s;
      ws[q][wid][1] = v[q].c;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = ((int)blockDim.x + 63) / 64;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      CSum a{ws[q][0][0], ws[q][0][1]};
      for (int w = 1; w < nw; ++w) a.merge(ws[q][w][0], ws[q][w][1]);
      partials[blk * NQ + q] = make_double2(a.s, a.c);
    }
  }
}

// Final deterministic reduction of nblk x nq partials -> out[q] = s + c, one
// block of 1024 threads per batch member (partials sp, outputs so apart): fixed strided assignment, 8 loads in flight per
// thread, xor butterflies in the wavefronts, wavefront results merged in order.
static __global__ __launch_bounds__(1024) void k_reduce_partials_wide(const double2* __restrict__ partials, long nblk,
                                                                      int nq, double* __restrict__ out, long sp,
                                                                      long so) {
  partials += blockIdx.x * sp;  // one block per batch member
  out += blockIdx.x * so;
  __shared__ double ws[16][2];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  for (int q = 0; q < nq; ++q) {
    CSum a{0.0, 0.0};
    for (long base = 0; base < nblk; base += 8 * 1024) {
      double2 v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const long i = base + u * 1024 + t;
        v[u] = i < nblk ? partials[i * nq + q] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) a.merge(v[u].x, v[u].y);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double s2 = __shfl_xor(a.s, off), c2 = __shfl_xor(a.c, off);
      a.merge(s2, c2);
    }
    if (lane == 0) {
      ws[wid][0] = a.s;
      ws[wid][1] = a.c;
    }
    __syncthreads();
    if (t == 0) {
      CSum r{ws[0][0], ws[0][1]};
      for (int w = 1; w < 16; ++w) r.merge(ws[w][0], ws[w][1]);
      out[q] = r.s + r.c;
    }
    __syncthreads();
  }
}

// Pass-1 registers of the inverse length-M transform of one half-spectrum row
// X_0..X_M (src points at X_0):
// z_j = (X_j + X*_{M-j}) + i W_N^{-j} (X_j - X*_{M-j}), j = R2 n1 + t + P q;
// the imaginary parts of X_0 and X_M are dropped (Hermitian input).  The lane
// part t is folded into two base pointers so that every load carries a
// constant immediate offset.
template <typename T, int M>
__device__ __forceinline__ void c2r_to_regs(const cplx<T>* __restrict__ src, const cplx<T>* __restrict__ twN,
                                            int t, cplx<T> (&v)[RowF<T, M>::Q1][RowF<T, M>::R1]) {
  using RF = RowF<T, M>;
  const cplx<T>* fw = src + t;        // X_{c + t}
  const cplx<T>* bw = src + (M - t);  // X_{M - t - c}
  const cplx<T>* tw = twN + t;
#pragma unroll
  for (int q = 0; q < RF::Q1; ++q)
#pragma unroll
    for (int n1 = 0; n1 < RF::R1; ++n1) {
      const int c = RF::R2 * n1 + RF::P * q;
      cplx<T> xk = fw[c];
      cplx<T> xm = bw[-c];
      if (n1 == 0 && q == 0 && t == 0) {
        xk.y = T(0);
        xm.y = T(0);
      }
      v[q][n1] = c2r_pre(xk, xm, cconj(tw[c]));
    }
}

// Half spectrum X_0..X_M of a row (written at dst) from the registers
// Z[k1 + R1 k2] (k1 = t + P q) of its forward length-M transform:
// X_k = (Z_k + Z*_{M-k})/2 - i W_N^k (Z_k - Z*_{M-k})/2,
// X_0 = Re Z_0 + Im Z_0, X_M = Re Z_0 - Im Z_0.  The caller has synchronised so
// that slot is free; every thread of the block must call this.
template <typename T, int M>
__device__ __forceinline__ void r2c_regs_to_slot(cplx<T>* slot, int t,
                                                 cplx<T> (&v)[RowF<T, M>::Q2][RowF<T, M>::R2]) {
  using RF = RowF<T, M>;
#pragma unroll
  for (int q = 0; q < RF::Q2; ++q)
#pragma unroll
    for (int k2 = 0; k2 < RF::R2; ++k2) slot[t + RF::P * q + RF::R1 * k2] = v[q][k2];
}

// Second half of r2c_from_regs (after the block has synchronised): the half spectrum from the slot to dst.
// Each thread takes the pairs (k, M - k), k <= M/2: with s = Z_k + Z*_{M-k}, d = Z_k - Z*_{M-k} and W^{M-k} =
// -conj(W^k), X_{M-k} = (s.x - (W^k d).y)/2 - i (s.y + (W^k d).x)/2 reuses the loads and the product of X_k.
template <typename T, int M>
__device__ __forceinline__ void r2c_slot_to_row(const cplx<T>* slot, int t, const cplx<T>* __restrict__ twN,
                                                cplx<T>* __restrict__ dst) {
  using RF = RowF<T, M>;
  constexpr int NI = (M / 2 + RF::P) / RF::P;
  // element M - k = M - t - P i sits at P (NI - 1 - i) past these bases (M - t - P (NI - 1) > 0 for every length),
  // so the mirrored loads and stores take constant offsets
  const cplx<T>* rs = slot + (M - t - RF::P * (NI - 1));
  cplx<T>* rd = dst + (M - t - RF::P * (NI - 1));
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int k = t + RF::P * i;
    if (k > M / 2) break;
    const bool k0 = i == 0 && t == 0;  // Z_0 holds X_0 and X_M
    const cplx<T> zk = slot[k];
    const cplx<T> zm = k0 ? zk : rs[RF::P * (NI - 1 - i)];
    const cplx<T> s = {zk.x + zm.x, zk.y - zm.y};  // Z_k + conj(Z_{M-k})
    const cplx<T> d = {zk.x - zm.x, zk.y + zm.y};  // Z_k - conj(Z_{M-k})
    const cplx<T> wd = cmul(twN[k], d);
    cplx<T> xk = {T(0.5) * (s.x + wd.y), T(0.5) * (s.y - wd.x)};
    cplx<T> xm = {T(0.5) * (s.x - wd.y), T(-0.5) * (s.y + wd.x)};
    if (k0) {
      xk = {zk.x + zk.y, T(0)};
      xm = {zk.x - zk.y, T(0)};
    }
    dst[k] = xk;
    if (2 * k != M) rd[RF::P * (NI - 1 - i)] = xm;
  }
}

template <typename T, int M>
__device__ __forceinline__ void r2c_from_regs(cplx<T>* slot, int t, const cplx<T>* __restrict__ twN,
                                              cplx<T>* __restrict__ dst, bool store,
                                              cplx<T> (&v)[RowF<T, M>::Q2][RowF<T, M>::R2]) {
  r2c_regs_to_slot<T, M>(slot, t, v);
  __syncthreads();
  if (store) r2c_slot_to_row<T, M>(slot, t, twN, dst);
}

// Real rows -> half spectra.  grid = (ceil(nrows/RB), narr), block = RB * P,
// dynamic LDS = RB * SLOT elements.  psum (optional): per-block compensated FP64 sums of the input rows of array
// (member) blockIdx.y at psum[blockIdx.y * psp + block].
template <typename T, typename TI, int M>
__global__ __launch_bounds__(256) void k_rows_r2c2(const TI* __restrict__ in, long in_arr_stride,
                                                   cplx<T>* __restrict__ out, long out_arr_stride, long nrows, int Zp, const cplx<T>* __restrict__ twP,
                                                   const cplx<T>* __restrict__ twN, int RB,
                                                   double2* __restrict__ psum = nullptr, long psp = 0,
                                                   Skip sk = Skip{}) {
  if (sk.off(blockIdx.y)) return;
  using RF = RowF<T, M>;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int seq = threadIdx.x / RF::P, t = threadIdx.x % RF::P;
  cplx<T>* slot = reinterpret_cast<cplx<T>*>(smem_raw) + seq * RF::SLOT;
  const long row0 = (long)row_tile<T>() * RB + seq;
  const bool ok = row0 < nrows;
  const long row = ok ? row0 : nrows - 1;
  const cplx<TI>* src = reinterpret_cast<const cplx<TI>*>(in + blockIdx.y * in_arr_stride + row * 2 * M) + t;
  cplx<T> v1[RF::Q1][RF::R1], v2[RF::Q2][RF::R2];
  CSum acc[1] = {{0.0, 0.0}};
#pragma unroll
  for (int q = 0; q < RF::Q1; ++q)
#pragma unroll
    for (int n1 = 0; n1 < RF::R1; ++n1) {
      const cplx<TI> x = src[RF::R2 * n1 + RF::P * q];
      v1[q][n1] = {T(x.x), T(x.y)};
      if (psum && ok) {
        acc[0].add(double(x.x));
        acc[0].add(double(x.y));
      }
    }
  reg_fft_seqp<T, M, RF::R1, RF::R2, true, 1>(v1, v2, t, slot, 1, twP);
  __syncthreads();
  r2c_from_regs<T, M>(slot, t, twN, out + blockIdx.y * out_arr_stride + row * Zp, ok, v2);
  if (psum) block_reduce_store_n<1>(acc, psum + blockIdx.y * psp, row_tile<T>());
}

// Half spectra -> real rows (unnormalised, times `scale`).
template <typename T, typename TO, int M>
__global__ __launch_bounds__(256) void k_rows_c2r2(const cplx<T>* __restrict__ in, long in_arr_stride,
                                                   TO* __restrict__ out, long out_arr_stride, long nrows, int Zp,
                                                   const cplx<T>* __restrict__ twP, const cplx<T>* __restrict__ twN,
                                                   int RB, double scale) {
  using RF = RowF<T, M>;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int seq = threadIdx.x / RF::P, t = threadIdx.x % RF::P;
  cplx<T>* slot = reinterpret_cast<cplx<T>*>(smem_raw) + seq * RF::SLOT;
  const long row0 = (long)row_tile<T>() * RB + seq;
  const bool ok = row0 < nrows;
  const long row = ok ? row0 : nrows - 1;
  cplx<T> v1[RF::Q1][RF::R1], v2[RF::Q2][RF::R2];
  c2r_to_regs<T, M>(in + blockIdx.y * in_arr_stride + row * Zp, twN, t, v1);
  reg_fft_seqp<T, M, RF::R1, RF::R2, false, 1>(v1, v2, t, slot, 1, twP);
  if (!ok) return;
  cplx<TO>* dst = reinterpret_cast<cplx<TO>*>(out + blockIdx.y * out_arr_stride + row * 2 * M) + t;
#pragma unroll
  for (int q = 0; q < RF::Q2; ++q)
#pragma unroll
    for (int k2 = 0; k2 < RF::R2; ++k2)
      dst[RF::P * q + RF::R1 * k2] = {TO(double(v2[q][k2].x) * scale), TO(double(v2[q][k2].y) * scale)};
}

// Pointwise z pass: inverse z of the nconv arrays of RB rows, Phi / dPhi at
// every site, forward z of the m dPhi rows back into B, v_mf = the (m+1)-th
// array in real space, F partial sums.  Sequence s = a * RB + r.
// grid = (ceil(nrows/RB), members), block = nconv * RB * P, dynamic LDS = nconv * RB * SLOT elements.
// trace != 0 (tensor models whose weights satisfy w_Txx + w_Tyy + w_Tzz = w_s): only m = 10 arrays are stored,
// n_Tzz = n_s - n_Txx - n_Tyy per site, and the T_zz derivative is folded back onto the arrays it was built
// from (w_Tzz * dPhi_Tzz = w_s * dPhi_Tzz - w_Txx * dPhi_Tzz - w_Tyy * dPhi_Tzz).
// ms (mixed path, null otherwise): the FP32 arrays hold the weighted densities without their means, which are
// added back per site in FP64 (ms[kMSn + a]); dPhi_a(nbar) (ms[kMSc + a]) is subtracted before the FP32 store.
template <typename T, int MODEL, int M>
__global__ __launch_bounds__(256, 3) void k_z_pointwise2(cplx<T>* __restrict__ B, long arr_stride, int m, int nconv,
                                                      double* __restrict__ vmf, long nrows, int Zp,
                                                      const cplx<T>* __restrict__ twP,
                                                      const cplx<T>* __restrict__ twN, int RB, PhiParams pp,
                                                      double2* __restrict__ partials, int* __restrict__ flag,
                                                      Batch bt, int trace, const double* __restrict__ ms) {
  if (bt.sk.off(blockIdx.y)) return;
  B += blockIdx.y * bt.sb;
  if (ms) ms += blockIdx.y * kMS;
  if (vmf) vmf += blockIdx.y * bt.sr;
  partials += blockIdx.y * bt.sp;
  flag += blockIdx.y;
  using RF = RowF<T, M>;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int seq = threadIdx.x / RF::P, t = threadIdx.x % RF::P;
  const int a = seq / RB, r = seq % RB;
  cplx<T>* slot = lds + seq * RF::SLOT;
  const long row0 = (long)row_tile<T>() * RB;
  const long rrow = row0 + r < nrows ? row0 + r : nrows - 1;
  const bool ok = row0 + r < nrows;
  cplx<T>* brow = B + a * arr_stride + rrow * Zp;
  // 1. inverse z of every array into the slots (natural order: complex m = sites 2m, 2m+1)
  {
    cplx<T> v1[RF::Q1][RF::R1], v2[RF::Q2][RF::R2];
    c2r_to_regs<T, M>(brow, twN, t, v1);
    reg_fft_seqp<T, M, RF::R1, RF::R2, false, 1>(v1, v2, t, slot, 1, twP);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RF::Q2; ++q)
#pragma unroll
      for (int k2 = 0; k2 < RF::R2; ++k2) slot[t + RF::P * q + RF::R1 * k2] = v2[q][k2];
  }
  __syncthreads();
  // 2. Phi and dPhi per site
  CSum F{0.0, 0.0};
  int bad = 0;
  const bool has_mf = nconv > m;
  for (int s = threadIdx.x; s < RB * 2 * M; s += blockDim.x) {
    const int rr = s / (2 * M), z = s % (2 * M);
    const long row = row0 + rr;
    if (row >= nrows) continue;
    double n[11];
#pragma unroll
    for (int b = 0; b < 11; ++b)
      n[b] = b < m ? double(reinterpret_cast<const T*>(lds + (b * RB + rr) * RF::SLOT)[z]) + (ms ? ms[kMSn + b] : 0.0)
                   : 0.0;
    if (trace) n[10] = n[1] - n[5] - n[8];
    double out[11];
    const double phi = phi_site<MODEL>(n, pp, out);
    if (trace) {
      out[1] += out[10];
      out[5] -= out[10];
      out[8] -= out[10];
    }
    if (n[0] > 0.5 && 1 - n[0] < 0.0) bad = 1;
    F.add(phi);
#pragma unroll
    for (int b = 0; b < 11; ++b)
      if (b < m) reinterpret_cast<T*>(lds + (b * RB + rr) * RF::SLOT)[z] = T(out[b] - (ms ? ms[kMSc + b] : 0.0));
    if (has_mf)
      vmf[row * 2 * M + z] = double(reinterpret_cast<const T*>(lds + (m * RB + rr) * RF::SLOT)[z]) +
                             (ms ? ms[kMSn + m] : 0.0);
  }
  if (bad) atomicOr(flag, 1);
  __syncthreads();
  // 3. forward z of the dPhi rows (the mean-field sequences run along, storing nothing)
  {
    cplx<T> v1[RF::Q1][RF::R1], v2[RF::Q2][RF::R2];
#pragma unroll
    for (int q = 0; q < RF::Q1; ++q)
#pragma unroll
      for (int n1 = 0; n1 < RF::R1; ++n1) v1[q][n1] = slot[RF::R2 * n1 + RF::P * q + t];
    __syncthreads();
    reg_fft_seqp<T, M, RF::R1, RF::R2, true, 1>(v1, v2, t, slot, 1, twP);
    __syncthreads();
    r2c_from_regs<T, M>(slot, t, twN, brow, ok && a < m, v2);
  }
  CSum v[1] = {F};
  block_reduce_store_n<1>(v, partials, row_tile<T>());
}

// Pointwise z pass of the Hessian-vector product (FP64): sequences a < m are the
// weighted densities n_a of rho (array a of B), m <= a < 2m the weighted
// densities psi_a of the direction v (array a - m of B2), a = 2m (mf) the
// mean-field convolution of v.  Per site: Lambda = e part of the collapsed dPhi
// at n + psi e (phi_site on dual numbers), written over n_a and transformed back
// into B; v_mf <- the mean-field row of v.  Reference: FMT::add_second_derivative
// (src/FMT.cpp:437-567), whose back-convolution with conjugated weights equals
// the parity-signed collapse used for dF.
template <int MODEL, int M>
__global__ __launch_bounds__(512) void k_z_hess2(cplx<double>* __restrict__ B, const cplx<double>* __restrict__ B2,
                                                 long arr_stride, int m, int mf, double* __restrict__ vmf, long nrows,
                                                 int Zp, const cplx<double>* __restrict__ twP,
                                                 const cplx<double>* __restrict__ twN, int RB, PhiParams pp,
                                                 double inv_dp, double scale, int* __restrict__ flag, Batch bt,
                                                 int trace, int single_offdiag) {
  B += blockIdx.y * bt.sb;
  B2 += blockIdx.y * bt.sb;
  if (vmf) vmf += blockIdx.y * bt.sr;
  flag += blockIdx.y;
  using RF = RowF<double, M>;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
  const int seq = threadIdx.x / RF::P, t = threadIdx.x % RF::P;
  const int a = seq / RB, r = seq % RB;
  cplx<double>* slot = lds + seq * RF::SLOT;
  const long row0 = (long)blockIdx.x * RB;
  const long rrow = row0 + r < nrows ? row0 + r : nrows - 1;
  const bool ok = row0 + r < nrows;
  cplx<double>* brow = B + (a < m ? a : 0) * arr_stride + rrow * Zp;
  const cplx<double>* src = a < m ? brow : B2 + (a - m) * arr_stride + rrow * Zp;
  {
    cplx<double> v1[RF::Q1][RF::R1], v2[RF::Q2][RF::R2];
    c2r_to_regs<double, M>(src, twN, t, v1);
    reg_fft_seqp<double, M, RF::R1, RF::R2, false, 1>(v1, v2, t, slot, 1, twP);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RF::Q2; ++q)
#pragma unroll
      for (int k2 = 0; k2 < RF::R2; ++k2) slot[t + RF::P * q + RF::R1 * k2] = v2[q][k2];
  }
  __syncthreads();
  int bad = 0;
  for (int s = threadIdx.x; s < RB * 2 * M; s += blockDim.x) {
    const int rr = s / (2 * M), z = s % (2 * M);
    const long row = row0 + rr;
    if (row >= nrows) continue;
    double nv[11], pv[11];
#pragma unroll
    for (int b = 0; b < 11; ++b) {
      nv[b] = b < m ? reinterpret_cast<const double*>(lds + (b * RB + rr) * RF::SLOT)[z] : 0.0;
      pv[b] = b < m ? reinterpret_cast<const double*>(lds + ((m + b) * RB + rr) * RF::SLOT)[z] : 0.0;
    }
    if (trace) {  // T_zz from the trace identity, as in k_z_pointwise2
      nv[10] = nv[1] - nv[5] - nv[8];
      pv[10] = pv[1] - pv[5] - pv[8];
    }
    // measures of rho (diameter 1/inv_d) carrying those of the direction (diameter 1/inv_dp) as e parts
    const double id = pp.inv_d, id2 = pp.inv_d2, ip2 = inv_dp * inv_dp;
    const Dual eta(nv[0], pv[0]), s0(nv[1] * id2, pv[1] * ip2), s1(nv[1] * id, pv[1] * inv_dp), s2(nv[1], pv[1]);
    const Dual v1[3] = {Dual(nv[2] * id, pv[2] * inv_dp), Dual(nv[3] * id, pv[3] * inv_dp),
                        Dual(nv[4] * id, pv[4] * inv_dp)};
    const Dual v2[3] = {Dual(nv[2], pv[2]), Dual(nv[3], pv[3]), Dual(nv[4], pv[4])};
    const Dual T6[6] = {Dual(nv[5], pv[5]), Dual(nv[6], pv[6]), Dual(nv[7], pv[7]),
                        Dual(nv[8], pv[8]), Dual(nv[9], pv[9]), Dual(nv[10], pv[10])};
    PhiDeriv<Dual> d;
    phi_meas<MODEL, Dual>(eta, s0, s1, s2, v1, v2, T6, pp, d);
    Dual out[11];
    phi_collapse<(MODEL >= M_ESFMT), Dual>(d, id, id2, out);
    if (trace) {
      out[1] = out[1] + out[10];
      out[5] = out[5] - out[10];
      out[8] = out[8] - out[10];
    }
    if (single_offdiag) {  // the reference's AO force: one of T_jk, T_kj per off-diagonal array (Species.h:406-425)
      out[6] = 0.5 * out[6];
      out[7] = 0.5 * out[7];
      out[9] = 0.5 * out[9];
    }
    if (nv[0] > 0.5 && 1 - nv[0] < 0.0) bad = 1;
#pragma unroll
    for (int b = 0; b < 11; ++b)
      if (b < m) reinterpret_cast<double*>(lds + (b * RB + rr) * RF::SLOT)[z] = scale * out[b].d;
    if (mf) vmf[row * 2 * M + z] = reinterpret_cast<const double*>(lds + (2 * m * RB + rr) * RF::SLOT)[z];
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
    r2c_from_regs<double, M>(slot, t, twN, brow, ok && a < m, v2);
  }
}

// Pointwise z pass of an FMT species with Asakura-Oosawa polymer (FP64): as k_z_pointwise2, and the same dPhi
// collapsed a second time with the polymer diameter 1/inv_dp into the m arrays of B2 (FMT_AO_Species::
// set_fundamental_measure_derivatives, src/FMT_Species.cpp:663-686).  Sequences a < nconv load array a of B;
// sequences nconv + a (a < m) only hold the AO rows (their inverse transform runs on zeros so that every thread
// meets the transform's barrier).
template <int MODEL, int M>
__global__ __launch_bounds__(512) void k_z_pw_ao(cplx<double>* __restrict__ B, cplx<double>* __restrict__ B2,
                                                 long arr_stride, int m, int nconv, double* __restrict__ vmf,
                                                 long nrows, int Zp, const cplx<double>* __restrict__ twP,
                                                 const cplx<double>* __restrict__ twN, int RB, PhiParams pp,
                                                 double inv_dp, double2* __restrict__ partials, int* __restrict__ flag,
                                                 Batch bt) {
  B += blockIdx.y * bt.sb;
  B2 += blockIdx.y * bt.sb;
  if (vmf) vmf += blockIdx.y * bt.sr;
  partials += blockIdx.y * bt.sp;
  flag += blockIdx.y;
  using RF = RowF<double, M>;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
  const int seq = threadIdx.x / RF::P, t = threadIdx.x % RF::P;
  const int a = seq / RB, r = seq % RB;
  cplx<double>* slot = lds + seq * RF::SLOT;
  const long row0 = (long)blockIdx.x * RB;
  const long rrow = row0 + r < nrows ? row0 + r : nrows - 1;
  const bool ok = row0 + r < nrows;
  const bool is_ao = a >= nconv;
  cplx<double>* brow = is_ao ? B2 + (a - nconv) * arr_stride + rrow * Zp : B + a * arr_stride + rrow * Zp;
  {
    cplx<double> v1[RF::Q1][RF::R1], v2[RF::Q2][RF::R2];
    if (!is_ao) {
      c2r_to_regs<double, M>(brow, twN, t, v1);
    } else {
#pragma unroll
      for (int q = 0; q < RF::Q1; ++q)
#pragma unroll
        for (int n1 = 0; n1 < RF::R1; ++n1) v1[q][n1] = {0.0, 0.0};
    }
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
  const bool has_mf = nconv > m;
  constexpr bool TENSOR = MODEL >= M_ESFMT;
  for (int s = threadIdx.x; s < RB * 2 * M; s += blockDim.x) {
    const int rr = s / (2 * M), z = s % (2 * M);
    const long row = row0 + rr;
    if (row >= nrows) continue;
    double n[11];
#pragma unroll
    for (int b = 0; b < 11; ++b) n[b] = b < m ? reinterpret_cast<const double*>(lds + (b * RB + rr) * RF::SLOT)[z] : 0.0;
    const double id = pp.inv_d, id2 = pp.inv_d2;
    const double v1[3] = {n[2] * id, n[3] * id, n[4] * id}, v2[3] = {n[2], n[3], n[4]};
    PhiDeriv<double> d;
    const double phi = phi_meas<MODEL, double>(n[0], n[1] * id2, n[1] * id, n[1], v1, v2, n + 5, pp, d);
    if (n[0] > 0.5 && 1 - n[0] < 0.0) bad = 1;
    F.add(phi);
    double out[11], outp[11];
    phi_collapse<TENSOR, double>(d, id, id2, out);
    phi_collapse<TENSOR, double>(d, inv_dp, inv_dp * inv_dp, outp);
#pragma unroll
    for (int b = 0; b < 11; ++b)
      if (b < m) {
        reinterpret_cast<double*>(lds + (b * RB + rr) * RF::SLOT)[z] = out[b];
        reinterpret_cast<double*>(lds + ((nconv + b) * RB + rr) * RF::SLOT)[z] = outp[b];
      }
    if (has_mf) vmf[row * 2 * M + z] = reinterpret_cast<const double*>(lds + (m * RB + rr) * RF::SLOT)[z];
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
    r2c_from_regs<double, M>(slot, t, twN, brow, ok && (a < m || is_ao), v2);
  }
  CSum v[1] = {F};
  block_reduce_store_n<1>(v, partials, blockIdx.x);
}

// AO free energy density and the field the polymer weights act on (FMT_AO_Species::free_energy_post_process,
// src/FMT_Species.cpp:690-738): psi <- exp(-psi) in place; partial sums of 1 - exp(-psi) per block.
// grid = (blocks, members), block = kThreads.
static __global__ __launch_bounds__(256) void k_ao_exp(double* __restrict__ psi, long n, long stride, double2* __restrict__ partials,
                                                long sp) {
  psi += blockIdx.y * stride;
  partials += blockIdx.y * sp;
  CSum acc{0.0, 0.0};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double e = exp(-psi[i]);
    psi[i] = e;
    acc.add(1.0 - e);
  }
  CSum v[1] = {acc};
  block_reduce_store_n<1>(v, partials, blockIdx.x);
}

// Final z pass: dF = [dF_in] + dV ln rho + dV c2r(acc) + v_mf + dV V_ext - dV mu
// (src/DFT.cpp:377-425), sites 2m, 2m+1 of each row straight from the
// registers of the inverse transform.  Hessian mode (fp.hess): out = [out_in] +
// dV v / rho (ideal) + dV c2r(acc) + v_mf, with v passed in the V_ext slot.  partials per block:
// {sum (rho ln rho - rho) dV, sum rho v_mf, sum V_ext rho, sum rho}.
template <typename T, int M>
__global__ __launch_bounds__(256) void k_z_final2(const cplx<T>* __restrict__ Acc, const double* __restrict__ rho,
                                                  const double* __restrict__ vext, const double* __restrict__ vmf,
                                                  double* __restrict__ dF, unsigned real_bytes, long nrows, int Zp,
                                                  const cplx<T>* __restrict__ twM, const cplx<T>* __restrict__ twN,
                                                  int RB, FinalParams fp, double2* __restrict__ partials, Batch bt) {
  if (bt.sk.off(blockIdx.y)) return;
  const long mb = blockIdx.y;  // batch member
  Acc += mb * bt.sa;
  rho += mb * bt.sr;
  dF += mb * bt.sr;
  if (vext) vext += mb * bt.sr;
  if (vmf) vmf += mb * bt.sr;
  partials += mb * bt.sp;
  if (bt.mu) fp.mu = bt.mu[mb];
  using RF = RowF<T, M>;
  // 8-byte-aligned view of the dynamic LDS and the plain twiddle table: 16-byte exchange accesses and the
  // pass-ordered table measured slower in this kernel (64-thread blocks) than in the other row kernels
  extern __shared__ unsigned char smem_fin[];
  const int seq = threadIdx.x / RF::P, t = threadIdx.x % RF::P;
  cplx<T>* slot = reinterpret_cast<cplx<T>*>(smem_fin) + seq * RF::SLOT;
  const long row0 = (long)row_tile<T>() * RB + seq;
  const bool ok = row0 < nrows;
  const long row = ok ? row0 : nrows - 1;
  cplx<T> y[RF::Q2][RF::R2];
  if (fp.fmt) {
    cplx<T> v1[RF::Q1][RF::R1];
    c2r_to_regs<T, M>(Acc + row * Zp, twN, t, v1);
    reg_fft_seq<T, M, RF::R1, RF::R2, false, 1>(v1, y, t, slot, 1, twM);
  } else {
#pragma unroll
    for (int q = 0; q < RF::Q2; ++q)
#pragma unroll
      for (int k2 = 0; k2 < RF::R2; ++k2) y[q][k2] = {T(0), T(0)};
  }
  CSum qs[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  const double fmt_mean = fp.ms ? fp.ms[mb * kMS + kMSk] : 0.0;
  if (ok) {
    const auto rr = buf_rsrc(rho, real_bytes);
    const auto rf = buf_rsrc(dF, real_bytes);
    const auto rm = buf_rsrc(vmf, real_bytes);
    const auto re = buf_rsrc(vext, real_bytes);
    const unsigned row_byte = unsigned(row * 2 * M) * 8u;
#pragma unroll
    for (int q = 0; q < RF::Q2; ++q)
#pragma unroll
      for (int k2 = 0; k2 < RF::R2; ++k2) {
        const unsigned off = row_byte + unsigned(t + RF::P * q + RF::R1 * k2) * 16u;
        const cplx<double> r2 = buf_ld<cplx<double>>(rr, off, 0);
        const cplx<double> f2 = fp.accumulate ? buf_ld<cplx<double>>(rf, off, 0) : cplx<double>{0.0, 0.0};
        const cplx<double> m2 = fp.mf ? buf_ld<cplx<double>>(rm, off, 0) : cplx<double>{0.0, 0.0};
        const cplx<double> e2 = fp.ext || fp.hess ? buf_ld<cplx<double>>(re, off, 0) : cplx<double>{0.0, 0.0};
        double out[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const double d0 = h ? r2.y : r2.x;
          const double vm = h ? m2.y : m2.x, ve = h ? e2.y : e2.x;
          double d = h ? f2.y : f2.x;
          if (fp.ideal && fp.hess) {
            d += ve / d0 * fp.dV;  // ideal-gas Hessian: V_ext slot carries the direction v
          } else if (fp.ideal) {
            const double l = log(d0);
            qs[0].add((d0 * l - d0) * fp.dV);
            d += l * fp.dV;
          }
          if (fp.fmt) d += double(h ? y[q][k2].y : y[q][k2].x) * fp.dV + fmt_mean;
          if (fp.mf) {
            d += vm;
            qs[1].add(d0 * vm);
          }
          if (fp.ext) {
            qs[2].add(ve * d0);
            d += ve * fp.dV;
          }
          qs[3].add(d0);
          d += -fp.mu * fp.dV;
          out[h] = d;
        }
        buf_st(rf, off, 0, cplx<double>{out[0], out[1]});
      }
  }
  block_reduce_store_n<4>(qs, partials, row_tile<T>());
}

}  // namespace cdft
