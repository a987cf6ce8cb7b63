// Pointwise z pass specialised on the number of stored arrays, the mean-field
// row and the rows per block (all compile-time): inverse z of the weighted
// densities, Phi / dPhi at every site in FP64, forward z of the dPhi rows.
// Same data layout, results and partial-sum slots as k_z_pointwise2.
//
// Compared with k_z_pointwise2:
// * the two twiddle tables of the row transforms are copied to LDS once per
//   block, so the transforms issue no vector-memory loads besides the rows;
//   the pass-1 table is in pass order (reg_fft_pass1p), read without bank
//   conflicts at constant offsets;
// * the per-site loads and stores of the n_alpha / dPhi_alpha values use
//   compile-time LDS offsets (no per-array address arithmetic or guards);
// * rows per block are chosen for the precision (three blocks of at most
//   53 KB of LDS per CU), and the per-site stage runs on four whole
//   wavefronts while the transforms use the P threads of each sequence.
#pragma once
#include "rows2.hpp"

namespace cdft {

template <typename T, int M, int NS>
struct Pw3Geo {
  using RF = RowF<T, M>;
  static constexpr int P = RF::P, SLOT = RF::SLOT;
  static constexpr int TW = 2 * M + 1;  // twP (M entries) + twN (entries 0..M)
  static constexpr int per_row = NS * SLOT * int(sizeof(cplx<T>));
  static constexpr int tw_bytes = TW * int(sizeof(cplx<T>));
  static constexpr int rows() {
    int rb = 1;
    while ((rb + 1) * NS * P <= 256 && (rb + 1) * per_row + tw_bytes <= 53 * 1024) ++rb;
    return rb;
  }
  static constexpr int RB = rows();
  static constexpr int NTF = NS * RB * P;                      // threads of the row transforms
  static constexpr int NT = NTF <= 256 ? 256 : (NTF + 63) / 64 * 64;  // threads per block (all of them in Phi)
  static constexpr int lds_bytes = RB * per_row + tw_bytes;
  static constexpr bool fits = lds_bytes <= 64 * 1024 && NT <= 1024;
};

// NA stored arrays (5 without tensor, 10 with T_zz from the trace identity, 11 otherwise), MF: one more sequence
// holding the mean-field convolution, written to vmf in real space.  ms: the mean split of the mixed path (as in
// k_z_pointwise2), null otherwise.  The row transforms run on the first NTF
// threads (P per sequence, sequence s = a RB + r for array a of row r); the per-site Phi stage spreads the sites
// of the block's rows over all NT threads (whole wavefronts).
// grid = (ceil(nrows / RB), members), block = NT, dynamic LDS = lds_bytes.
template <typename T, int MODEL, int M, int NA, bool MF>
__global__ __launch_bounds__((Pw3Geo<T, M, NA + (MF ? 1 : 0)>::NT)) void k_z_pw3(
    cplx<T>* __restrict__ B, long arr_stride, double* __restrict__ vmf, long nrows, int Zp,
    const cplx<T>* __restrict__ twP_g, const cplx<T>* __restrict__ twN_g, PhiParams pp,
    double2* __restrict__ partials, int* __restrict__ flag, Batch bt, const double* __restrict__ ms) {
  constexpr int NS = NA + (MF ? 1 : 0);
  using G = Pw3Geo<T, M, NS>;
  using RF = RowF<T, M>;
  constexpr int RB = G::RB, NT = G::NT, SLOT = RF::SLOT;
  constexpr bool TRACE = MODEL >= M_ESFMT && NA == 10;
  if (bt.sk.off(blockIdx.y)) return;
  B += blockIdx.y * bt.sb;
  if (ms) ms += blockIdx.y * kMS;
  if (MF) vmf += blockIdx.y * bt.sr;
  partials += blockIdx.y * bt.sp;
  flag += blockIdx.y;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* twP = lds + NS * RB * SLOT;
  cplx<T>* twN = twP + M;
  for (int i = threadIdx.x; i < G::TW; i += NT) twP[i] = i < M ? twP_g[i] : twN_g[i - M];
  const bool act = threadIdx.x < G::NTF;  // a transform thread
  const int seq = act ? threadIdx.x / RF::P : 0, t = threadIdx.x % RF::P;
  const int a = seq / RB, r = seq % RB;
  cplx<T>* slot = lds + seq * SLOT;
  const long row0 = (long)row_tile<T>() * RB;
  const long rrow = row0 + r < nrows ? row0 + r : nrows - 1;
  const bool ok = act && row0 + r < nrows;
  cplx<T>* brow = B + a * arr_stride + rrow * Zp;
  // 1. inverse z of every array into the slots (natural order: complex m = sites 2m, 2m+1)
  {
    cplx<T> v1[RF::Q1][RF::R1], v2[RF::Q2][RF::R2];
    __syncthreads();  // twiddle tables in place
    if (act) {
      c2r_to_regs<T, M>(brow, twN, t, v1);
      reg_fft_pass1p<T, M, RF::R1, RF::R2, false, 1>(v1, t, slot, 1, twP);
    }
    __syncthreads();
    if (act) reg_fft_pass2<T, M, RF::R1, RF::R2, false, 1>(v2, t, slot, 1);
    __syncthreads();
    if (act) {
#pragma unroll
      for (int q = 0; q < RF::Q2; ++q)
#pragma unroll
        for (int k2 = 0; k2 < RF::R2; ++k2) slot[t + RF::P * q + RF::R1 * k2] = v2[q][k2];
    }
  }
  __syncthreads();
  // 2. Phi and dPhi per site; array b of row rr sits at lds + (b RB + rr) SLOT
  CSum F{0.0, 0.0};
  int bad = 0;
  constexpr int NSITE = RB * 2 * M, NIT = (NSITE + NT - 1) / NT;
  constexpr int AST = RB * SLOT * 2;  // T elements between the slots of consecutive arrays of one row
  // the mean split's constants of this member, read once (uniform across the block)
  double nbar[NS], cbar[NA];
#pragma unroll
  for (int b = 0; b < NS; ++b) nbar[b] = ms ? ms[kMSn + b] : 0.0;
#pragma unroll
  for (int b = 0; b < NA; ++b) cbar[b] = ms ? ms[kMSc + b] : 0.0;
#pragma unroll 1
  for (int it = 0; it < NIT; ++it) {
    const int s = threadIdx.x + it * NT;
    if (s >= NSITE) break;
    const int rr = s / (2 * M), z = s % (2 * M);
    const long row = row0 + rr;
    if (row >= nrows) break;  // rows past the end come last in a thread's sequence of sites
    T* base = reinterpret_cast<T*>(lds + rr * SLOT) + z;
    double n[11];
#pragma unroll
    for (int b = 0; b < 11; ++b) n[b] = b < NA ? double(base[b * AST]) + nbar[b < NA ? b : 0] : 0.0;
    if (TRACE) n[10] = n[1] - n[5] - n[8];
    double out[11];
    const double phi = phi_site<MODEL>(n, pp, out);
    if (TRACE) {
      out[1] += out[10];
      out[5] -= out[10];
      out[8] -= out[10];
    }
    if (n[0] > 0.5 && 1 - n[0] < 0.0) bad = 1;
    F.add(phi);
#pragma unroll
    for (int b = 0; b < NA; ++b) base[b * AST] = T(out[b] - cbar[b]);
    if (MF) vmf[row * 2 * M + z] = double(base[NA * AST]) + nbar[NS - 1];
  }
  if (bad) atomicOr(flag, 1);
  __syncthreads();
  // 3. forward z of the dPhi rows (the mean-field sequences run along, storing nothing)
  {
    cplx<T> v1[RF::Q1][RF::R1], v2[RF::Q2][RF::R2];
    if (act) {
#pragma unroll
      for (int q = 0; q < RF::Q1; ++q)
#pragma unroll
        for (int n1 = 0; n1 < RF::R1; ++n1) v1[q][n1] = slot[RF::R2 * n1 + RF::P * q + t];
    }
    __syncthreads();
    if (act) reg_fft_pass1p<T, M, RF::R1, RF::R2, true, 1>(v1, t, slot, 1, twP);
    __syncthreads();
    if (act) reg_fft_pass2<T, M, RF::R1, RF::R2, true, 1>(v2, t, slot, 1);
    __syncthreads();
    if (act) r2c_regs_to_slot<T, M>(slot, t, v2);
    __syncthreads();
    if (ok && a < NA) r2c_slot_to_row<T, M>(slot, t, twN, brow);
  }
  CSum v[1] = {F};
  block_reduce_store_n<1>(v, partials, row_tile<T>());
}

// Host-side arguments of one pointwise launch (pointers already offset to the first x-plane of the pass).
struct PwArgs {
  void* B;
  long arr_stride;
  int m, nconv;
  double* vmf;
  long nrows;
  int Zp;
  const void* twP;  // pass-1 twiddles of length M in pass order (upload_pass_twiddles)
  const void* twN;
  PhiParams pp;
  double2* partials;
  int* flag;
  Batch bt;
  int trace;
  int model;
  int members;
  hipStream_t stream;
  const double* ms;
};

}  // namespace cdft
