// Column-pass kernels (x and y transforms) built on the register FFT of
// fft2.hpp.  Same data layout and results as k_cols / k_conv_fwd / k_conv_bwd
// in kernels.hpp.
//
// Columns are numbered flat: column c lives at element (c / cz) * os + c % cz,
// its j-th element es further on.  For the x pass cz = Ny * Zp (every (y, kz)
// pair, padding included), for the y pass cz = Zp, os = Ny * Zp.  A block
// handles NB consecutive columns with P = gcd(R1, R2) threads each (lane
// layout: column fastest), so small grids still fill whole wavefronts and the
// loads of one wavefront cover NB contiguous elements per row.
//
// Addressing: every array is read and written through a buffer resource;
// the part of an offset common to the block (row of the transform) goes in the
// scalar offset, the per-lane part in one 32-bit vector offset per sequence,
// so a thread keeps one offset register per sequence instead of a 64-bit
// address per element.  Arrays are below 2 GiB (checked by the host).
#pragma once
#include "fft2.hpp"
#include "kernels.hpp"

namespace cdft {

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

template <typename U>
__device__ __forceinline__ U buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(U) == 16)
    return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
  else if constexpr (sizeof(U) == 8)
    return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
  else
    return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}

template <typename U>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, U v) {
  if constexpr (sizeof(U) == 16) {
    using V = __attribute__((ext_vector_type(4))) unsigned;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V, v), r, (int)voff, (int)soff, 0);
  } else if constexpr (sizeof(U) == 8) {
    using V = __attribute__((ext_vector_type(2))) unsigned;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), r, (int)voff, (int)soff, 0);
  } else {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 0);
  }
}

// Members of a batch that a FIRE batch has stopped (converged, or unable to go on): every block of such a member
// returns at once, so a step costs what its running members need.  flag null: every member runs.
struct Skip {
  const int* flag = nullptr;  // member m runs when flag[m * stride] != 0
  long stride = 0;
  __device__ __forceinline__ bool off(unsigned m) const { return flag && !flag[m * stride]; }
};

// Member strides of a batch of same-shape systems (member = blockIdx.y of the
// kernels below): one complex array, the block of per-weight complex arrays, a
// real array, the partial-sum slots; per-member chemical potentials.
struct Batch {
  long sa, sb, sr, sp;
  const double* mu;  // null: use the scalar of the call
  Skip sk{};
};

// Logical tile of this block such that consecutive tiles run on one XCD.  The hardware hands the blocks of a
// launch to the 8 XCDs round-robin (block b to XCD b % 8), each XCD with its own L2; tiles that share cache lines
// (neighbouring column groups, rows that straddle a line) then fetch every line once per XCD.  Giving XCD k the
// contiguous tile range [k per + min(k, rem), ...) keeps neighbours on one L2.  A bijection of [0, gridDim.x).
__device__ __forceinline__ unsigned xcd_tile() {
  const unsigned n = gridDim.x, b = blockIdx.x;
  const unsigned per = n >> 3, rem = n & 7u, k = b & 7u;
  return k * per + (k < rem ? k : rem) + (b >> 3);
}

// Row kernels: FP32 half-spectrum rows (Zp x 8 bytes) straddle cache lines, so neighbouring rows go to one XCD;
// FP64 rows are whole lines and keep the hardware order (measured faster for the pointwise pass).
template <typename T>
__device__ __forceinline__ unsigned row_tile() {
  return sizeof(T) == 4 ? xcd_tile() : blockIdx.x;
}

struct ColGeo {
  unsigned bytes;  // bytes of one complex array (Nx * Ny * Zp elements)
  unsigned ncols;  // number of columns
  unsigned cz;     // columns per contiguous group
  unsigned os;     // element stride between groups
  unsigned es;     // element stride along the transform
  // Folded weight spectra (x passes of the convolutions; wq = 0: the full Nx x Ny x Zp layout).  Every weight is
  // even or odd in x (pxm: bit a set when array a is odd), so W(-kx, ky, kz) = +-W(kx, ky, kz): only the rows
  // kx <= Nx/2 are stored, (Nx/2+1) x Ny x Zp values per array, wstride apart.  A column reads its rows kx and
  // Nx - kx from the same stored row, so the fold also halves the distinct lines each block fetches.
  unsigned wq, wstride, pxm;
  Skip sk;  // stopped batch members (member = blockIdx.z of k_cols2)
};

// In-place transform of ncols columns per array.  grid = (ceil(ncols/NB), narr, members), block = NB * P.
template <typename T, int L, int NB, bool FWD>
__global__ __launch_bounds__(256) void k_cols2(cplx<T>* __restrict__ data, long arr_stride, ColGeo cg,
                                               const cplx<T>* __restrict__ tw, long mstride) {
  if (cg.sk.off(blockIdx.z)) return;
  data += blockIdx.z * mstride;  // batch member
  using F = F2<L>;
  constexpr int R1 = F::R1, R2 = F::R2, P = F::P, Q1 = F::Q1, Q2 = F::Q2;
  __shared__ cplx<T> buf[L * NB];
  const int b = threadIdx.x % NB, t = threadIdx.x / NB;
  const unsigned c0 = xcd_tile() * NB + b;
  const bool ok = c0 < cg.ncols;
  const unsigned c = ok ? c0 : cg.ncols - 1;  // tail lanes recompute the last column, store nothing
  const auto rd = buf_rsrc(data + blockIdx.y * arr_stride, cg.bytes);
  const unsigned col = (c / cg.cz) * cg.os + c % cg.cz;
  constexpr unsigned E = sizeof(cplx<T>);
  cplx<T> v1[Q1][R1], v2[Q2][R2];
#pragma unroll
  for (int q = 0; q < Q1; ++q) {
    const unsigned vo = (col + (t + P * q) * cg.es) * E;
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1)
      v1[q][n1] = buf_ld<cplx<T>>(rd, vo, R2 * n1 * cg.es * E);
  }
  reg_fft_seq<T, L, R1, R2, FWD>(v1, v2, t, buf + b, NB, tw);
  if (!ok) return;
#pragma unroll
  for (int q = 0; q < Q2; ++q) {
    const unsigned vo = (col + (t + P * q) * cg.es) * E;
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) buf_st(rd, vo, R1 * k2 * cg.es * E, v2[q][k2]);
  }
}

// x pass of rho_k, then per weight a: n_a(k) = rho_k w_a(k) and its inverse x
// pass, straight from registers.  grid = (ceil(ncols/NB), members), block = NB * P.
// zero_mean: rho_k(0) is dropped (the mean of every n_a is added back in FP64 by the pointwise kernel).
// ACC: the products are added to B instead of written (a mixture's measures summed over its species).
template <typename T, int L, int NB, bool ACC = false>
__global__ __launch_bounds__(256) void k_conv_fwd2(const cplx<T>* __restrict__ A, cplx<T>* __restrict__ B,
                                                   long arr_stride, const T* __restrict__ W, int nconv,
                                                   unsigned odd_mask, ColGeo cg, const cplx<T>* __restrict__ tw,
                                                   Batch bt, int zero_mean) {
  if (bt.sk.off(blockIdx.y)) return;
  A += blockIdx.y * bt.sa;
  B += blockIdx.y * bt.sb;
  using F = F2<L>;
  constexpr int R1 = F::R1, R2 = F::R2, P = F::P, Q1 = F::Q1, Q2 = F::Q2;
  __shared__ cplx<T> buf[L * NB];
  const int b = threadIdx.x % NB, t = threadIdx.x / NB;
  const unsigned c0 = xcd_tile() * NB + b;
  const bool ok = c0 < cg.ncols;
  const unsigned c = ok ? c0 : cg.ncols - 1;  // tail lanes recompute the last column, store nothing
  constexpr unsigned E = sizeof(cplx<T>), EW = sizeof(T);
  const auto ra = buf_rsrc(A, cg.bytes);
  cplx<T> v1[Q1][R1], rk[Q2][R2];
#pragma unroll
  for (int q = 0; q < Q1; ++q) {
    const unsigned vo = (c + (t + P * q) * cg.es) * E;
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1)
      v1[q][n1] = buf_ld<cplx<T>>(ra, vo, R2 * n1 * cg.es * E);
  }
  auto load_w = [&](int a, T (&w)[Q2][R2]) {
    if (cg.wq) {
      const auto rw = buf_rsrc(W + (long)a * cg.wstride, cg.wstride * EW);
#pragma unroll
      for (int q = 0; q < Q2; ++q)
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) {
          const unsigned kx = t + P * q + R1 * k2;
          const bool fx = kx > (unsigned)L / 2;
          const T v = buf_ld<T>(rw, (c + (fx ? L - kx : kx) * cg.es) * EW, 0);
          w[q][k2] = (fx && ((cg.pxm >> a) & 1u)) ? -v : v;
        }
      return;
    }
    const auto rw = buf_rsrc(W + a * arr_stride, cg.bytes / 2);
#pragma unroll
    for (int q = 0; q < Q2; ++q) {
      const unsigned vo = (c + (t + P * q) * cg.es) * EW;
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) w[q][k2] = buf_ld<T>(rw, vo, R1 * k2 * cg.es * EW);
    }
  };
  reg_fft_seq<T, L, R1, R2, true>(v1, rk, t, buf + b, NB, tw);
  if (zero_mean && c0 == 0 && t == 0) rk[0][0] = {T(0), T(0)};  // column (ky, kz) = 0, row kx = 0
  for (int a = 0; a < nconv; ++a) {
    const auto rb = buf_rsrc(B + a * arr_stride, cg.bytes);
    const bool odd = (odd_mask >> a) & 1u;
    cplx<T> u1[Q2][R2], u2[Q1][R1];
    T w[Q2][R2];
    load_w(a, w);
#pragma unroll
    for (int q = 0; q < Q2; ++q)
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) u1[q][k2] = wmul(rk[q][k2], w[q][k2], odd);
    // Re-read the twiddles in every iteration: kept live across the loop they
    // would cost 4 (L/P - 1) registers and halve the occupancy.
    int zero = 0;
    asm volatile("" : "+s"(zero));
    const cplx<T>* twa = tw + zero;
    __syncthreads();  // the exchange slot is reused
    reg_fft_seq<T, L, R2, R1, false>(u1, u2, t, buf + b, NB, twa);
    if (ok) {
#pragma unroll
      for (int q = 0; q < Q1; ++q) {
        const unsigned vo = (c + (t + P * q) * cg.es) * E;
#pragma unroll
        for (int k2 = 0; k2 < R1; ++k2) {
          cplx<T> o = u2[q][k2];
          if (ACC) {
            const cplx<T> prev = buf_ld<cplx<T>>(rb, vo, R2 * k2 * cg.es * E);
            o.x += prev.x;
            o.y += prev.y;
          }
          buf_st(rb, vo, R2 * k2 * cg.es * E, o);
        }
      }
    }
  }
}

// x pass of each dPhi_a, acc = sum_a w_a(k) dPhi_a(k) in registers, inverse x
// pass of acc.  grid = (ceil(ncols/NB), members), block = NB * P.  The weights of array a
// are loaded as one batch after its transform (all loads in flight at once).
template <typename T, int L, int NB>
__global__ __launch_bounds__(256) void k_conv_bwd2(const cplx<T>* __restrict__ D, cplx<T>* __restrict__ Acc,
                                                   long arr_stride, const T* __restrict__ W, int m,
                                                   unsigned odd_mask, ColGeo cg, const cplx<T>* __restrict__ tw,
                                                   Batch bt) {
  if (bt.sk.off(blockIdx.y)) return;
  D += blockIdx.y * bt.sb;
  Acc += blockIdx.y * bt.sa;
  using F = F2<L>;
  constexpr int R1 = F::R1, R2 = F::R2, P = F::P, Q1 = F::Q1, Q2 = F::Q2;
  __shared__ cplx<T> buf[L * NB];
  const int b = threadIdx.x % NB, t = threadIdx.x / NB;
  const unsigned c0 = xcd_tile() * NB + b;
  const bool ok = c0 < cg.ncols;
  const unsigned c = ok ? c0 : cg.ncols - 1;  // tail lanes recompute the last column, store nothing
  constexpr unsigned E = sizeof(cplx<T>), EW = sizeof(T);
  auto load_d = [&](int a, cplx<T> (&v)[Q1][R1]) {
    const auto rd = buf_rsrc(D + a * arr_stride, cg.bytes);
#pragma unroll
    for (int q = 0; q < Q1; ++q) {
      const unsigned vo = (c + (t + P * q) * cg.es) * E;
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) v[q][n1] = buf_ld<cplx<T>>(rd, vo, R2 * n1 * cg.es * E);
    }
  };
  auto load_w = [&](int a, T (&w)[Q2][R2]) {
    if (cg.wq) {
      const auto rw = buf_rsrc(W + (long)a * cg.wstride, cg.wstride * EW);
#pragma unroll
      for (int q = 0; q < Q2; ++q)
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) {
          const unsigned kx = t + P * q + R1 * k2;
          const bool fx = kx > (unsigned)L / 2;
          const T v = buf_ld<T>(rw, (c + (fx ? L - kx : kx) * cg.es) * EW, 0);
          w[q][k2] = (fx && ((cg.pxm >> a) & 1u)) ? -v : v;
        }
      return;
    }
    const auto rw = buf_rsrc(W + a * arr_stride, cg.bytes / 2);
#pragma unroll
    for (int q = 0; q < Q2; ++q) {
      const unsigned vo = (c + (t + P * q) * cg.es) * EW;
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) w[q][k2] = buf_ld<T>(rw, vo, R1 * k2 * cg.es * EW);
    }
  };
  cplx<T> acc[Q2][R2];
#pragma unroll
  for (int q = 0; q < Q2; ++q)
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) acc[q][k2] = {T(0), T(0)};
  for (int a = 0; a < m; ++a) {
    const bool odd = (odd_mask >> a) & 1u;
    cplx<T> v1[Q1][R1], v2[Q2][R2];
    T w[Q2][R2];
    load_d(a, v1);
    int zero = 0;  // re-read the twiddles per iteration (register pressure)
    asm volatile("" : "+s"(zero));
    const cplx<T>* twa = tw + zero;
    if (a) __syncthreads();  // the exchange slot is reused
    reg_fft_seq<T, L, R1, R2, true>(v1, v2, t, buf + b, NB, twa);
    load_w(a, w);
#pragma unroll
    for (int q = 0; q < Q2; ++q)
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) {
        const cplx<T> p = wmul(v2[q][k2], w[q][k2], odd);
        acc[q][k2] = {acc[q][k2].x + p.x, acc[q][k2].y + p.y};
      }
  }
  cplx<T> u2[Q1][R1];
  __syncthreads();
  reg_fft_seq<T, L, R2, R1, false>(acc, u2, t, buf + b, NB, tw);
  if (!ok) return;
  const auto racc = buf_rsrc(Acc, cg.bytes);
#pragma unroll
  for (int q = 0; q < Q1; ++q) {
    const unsigned vo = (c + (t + P * q) * cg.es) * E;
#pragma unroll
    for (int k2 = 0; k2 < R1; ++k2) buf_st(racc, vo, R2 * k2 * cg.es * E, u2[q][k2]);
  }
}

}  // namespace cdft
