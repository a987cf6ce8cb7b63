// Hot-path kernels of one F + dF evaluation.
//
// Spectral layout in HBM: complex S[x][y][kz] with kz in [0, Zp), Zp = Nz/2+1
// padded to a multiple of 8 elements so every row starts on a 128-byte line;
// only kz < Zh = Nz/2+1 is ever read or written.  Real arrays keep the
// reference layout [x][y][z] (include/Lattice.h:120-126).
//
// One evaluation (m weighted densities, + 1 if the mean field is on) runs:
//   k_rows_r2c      rho rows -> z-spectrum                     (FFTW r2c, z pass)
//   k_cols<fwd>     y pass
//   k_conv_fwd      x pass, then per weight a: * w_a(k), inverse x pass   (m outputs)
//   k_cols<inv>     y pass on the m arrays
//   k_z_pointwise   per row: inverse z (c2r) of the m arrays, Phi / dPhi at each
//                   site, forward z (r2c) of the m dPhi rows (n_a, dPhi_a never
//                   touch HBM as real arrays); F partial sums
//   k_cols<fwd>     y pass on the m dPhi arrays
//   k_conv_bwd      x pass per array, accumulate sum_a w_a(k) dPhi_a(k),
//                   inverse x pass of the sum
//   k_cols<inv>     y pass
//   k_z_final       inverse z of the sum, x dV, plus ideal gas, mean field,
//                   external field and -mu (src/DFT.cpp:377-425)
// Reference: FMT::calculateFreeEnergyAndDerivatives (src/FMT.cpp:363-427),
// FMT_Weighted_Density::convoluteWith / add_weight_schur_dPhi_to_arg
// (include/FMT_Weighted_Density.h:41-54).
#pragma once
#include <hip/hip_runtime.h>

#include "fft.hpp"
#include "phi.hpp"

namespace cdft {

constexpr int kThreads = 256;

struct Geo {
  int Nx, Ny, Nz, Zh, Zp;
};

__device__ __forceinline__ long sidx(const Geo& g, int x, int y, int kz) { return ((long)x * g.Ny + y) * g.Zp + kz; }

// Compensated (Neumaier) accumulator (include/Summation.h:15-56) and the
// error-free combination of two of them.
// Compensated sums need their additions unfused: a caller's multiply contracted into t = s + v would make the
// recovered error term refer to a different sum.
#ifndef CDFT_EXACT_FP
#define CDFT_EXACT_FP _Pragma("clang fp contract(off)")
#endif

struct CSum {
  double s, c;
  __device__ __forceinline__ void add(double v) {
    CDFT_EXACT_FP
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
  }
  __device__ __forceinline__ void merge(double s2, double c2) {
    CDFT_EXACT_FP
    const double t = s + s2;
    const double bp = t - s;
    const double e = (s - (t - bp)) + (s2 - bp);  // TwoSum error
    s = t;
    c += c2 + e;
  }
};

// Runs use(i, load(i)) for i in [0, n) across the workgroup with U independent
// global loads in flight per thread before any of them is consumed: the
// global -> LDS phases are latency-bound unless many loads overlap.
template <int U, typename V, typename Load, typename Use>
__device__ __forceinline__ void batched(int n, Load load, Use use) {
  for (int base = 0; base < n; base += U * kThreads) {
    V v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * kThreads + (int)threadIdx.x;
      if (i < n) v[u] = load(i);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * kThreads + (int)threadIdx.x;
      if (i < n) use(i, v[u]);
    }
  }
}

template <typename T>
struct cpair {
  cplx<T> a, b;
};

// Block reduction of NQ compensated sums; thread 0 writes partials[blk*NQ+q].
template <int NQ>
__device__ void block_reduce_store(CSum* v, double2* __restrict__ partials, long blk) {
  __shared__ double rs[NQ][kThreads], rc[NQ][kThreads];
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < NQ; ++q) { rs[q][t] = v[q].s; rc[q][t] = v[q].c; }
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (t < off) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        CSum a{rs[q][t], rc[q][t]};
        a.merge(rs[q][t + off], rc[q][t + off]);
        rs[q][t] = a.s;
        rc[q][t] = a.c;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) partials[blk * NQ + q] = make_double2(rs[q][0], rc[q][0]);
  }
}

// Final deterministic reduction of nblk x nq partials -> out[q] (one block).
static __global__ void k_reduce_partials(const double2* __restrict__ partials, long nblk, int nq, double* __restrict__ out) {
  __shared__ double rs[kThreads], rc[kThreads];
  const int t = threadIdx.x;
  for (int q = 0; q < nq; ++q) {
    CSum a{0.0, 0.0};
    batched<16, double2>(
        (int)nblk, [&](int b) { return partials[(long)b * nq + q]; }, [&](int, double2 p) { a.merge(p.x, p.y); });
    rs[t] = a.s;
    rc[t] = a.c;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
      if (t < off) {
        CSum b{rs[t], rc[t]};
        b.merge(rs[t + off], rc[t + off]);
        rs[t] = b.s;
        rc[t] = b.c;
      }
      __syncthreads();
    }
    if (t == 0) out[q] = rs[0] + rc[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Real <-> half-spectrum along z, via one complex FFT of length M = Nz/2:
// z_m = x_{2m} + i x_{2m+1};  X_k = (Z_k + Z*_{M-k})/2 - i W^k (Z_k - Z*_{M-k})/2,
// W = exp(-2 pi i / Nz), and the inverse
// z_k = (X_k + X*_{M-k}) + i W^{-k} (X_k - X*_{M-k}), IFFT_M(z)_m = x_{2m} + i x_{2m+1}.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void r2c_post(cplx<T> Zk, cplx<T> Zmk, cplx<T> w, cplx<T>& Xk) {
  const cplx<T> s = {Zk.x + Zmk.x, Zk.y - Zmk.y};  // Zk + conj(Zmk)
  const cplx<T> d = {Zk.x - Zmk.x, Zk.y + Zmk.y};  // Zk - conj(Zmk)
  const cplx<T> wd = cmul(w, d);
  // X = s/2 - i wd/2 = (s.x + wd.y)/2 + i (s.y - wd.x)/2
  Xk = {T(0.5) * (s.x + wd.y), T(0.5) * (s.y - wd.x)};
}

template <typename T>
__device__ __forceinline__ cplx<T> c2r_pre(cplx<T> Xk, cplx<T> Xmk, cplx<T> wconj) {
  const cplx<T> s = {Xk.x + Xmk.x, Xk.y - Xmk.y};
  const cplx<T> d = {Xk.x - Xmk.x, Xk.y + Xmk.y};
  const cplx<T> wd = cmul(wconj, d);
  // z = s + i wd
  return {s.x - wd.y, s.y + wd.x};
}

// Load one half-spectrum row (length M+1 in HBM) and write z_k (k < M) into lds row.
template <typename T>
__device__ __forceinline__ void load_row_c2r(const cplx<T>* __restrict__ src, cplx<T>* row, int M,
                                             const cplx<T>* __restrict__ twN, int k) {
  // k in [0, M/2]; handles the pair (k, M-k)
  cplx<T> Xk = src[k], Xmk = src[M - k];
  if (k == 0) {  // imaginary parts of X_0 and X_M are dropped (Hermitian input)
    Xk.y = T(0);
    Xmk.y = T(0);
  }
  row[k] = c2r_pre(Xk, Xmk, cconj(twN[k]));
  if (k != 0 && 2 * k != M) row[M - k] = c2r_pre(Xmk, Xk, cconj(twN[M - k]));
}

// Batched form of load_row_c2r: the load phase fetches X_k, X_{M-k} and both
// twiddles, the use phase writes z_k, z_{M-k} into the LDS row.
template <typename T>
struct c2r_item {
  cplx<T> xk, xmk, wk, wmk;
};

template <typename T>
__device__ __forceinline__ c2r_item<T> fetch_c2r(const cplx<T>* __restrict__ src, int M,
                                                 const cplx<T>* __restrict__ twN, int k) {
  return {src[k], src[M - k], twN[k], twN[M - k]};
}

template <typename T>
__device__ __forceinline__ void put_c2r(cplx<T>* row, int M, int k, c2r_item<T> it) {
  if (k == 0) {
    it.xk.y = T(0);
    it.xmk.y = T(0);
  }
  row[k] = c2r_pre(it.xk, it.xmk, cconj(it.wk));
  if (k != 0 && 2 * k != M) row[M - k] = c2r_pre(it.xmk, it.xk, cconj(it.wmk));
}

template <typename T>
__device__ __forceinline__ void zero_c2r(cplx<T>* row, int M, int k) {
  row[k] = {T(0), T(0)};
  if (k != 0 && 2 * k != M) row[M - k] = {T(0), T(0)};
}

template <typename T>
__device__ __forceinline__ void store_row_r2c(const cplx<T>* row, cplx<T>* __restrict__ dst, int M,
                                              const cplx<T>* __restrict__ twN, int k) {
  const cplx<T> Zk = row[k];
  const cplx<T> Zmk = row[k == 0 ? 0 : M - k];
  if (k == 0) {
    dst[0] = {Zk.x + Zk.y, T(0)};
    dst[M] = {Zk.x - Zk.y, T(0)};
    return;
  }
  cplx<T> X;
  r2c_post(Zk, Zmk, twN[k], X);
  dst[k] = X;
  if (2 * k != M) {
    r2c_post(Zmk, Zk, twN[M - k], X);
    dst[M - k] = X;
  }
}

// ---------------------------------------------------------------------------
// z pass, real -> spectrum, for narr arrays (blockIdx.y) of Nx*Ny rows.
// RB rows per block; LDS holds RB sequences of length M.
// ---------------------------------------------------------------------------
template <typename T, typename TI, int MAXPTS>
__global__ __launch_bounds__(kThreads) void k_rows_r2c(const TI* __restrict__ in, long in_arr_stride,
                                                       cplx<T>* __restrict__ out, long out_arr_stride, Geo g,
                                                       Radices radM, const cplx<T>* __restrict__ twM,
                                                       const cplx<T>* __restrict__ twN, int RB) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  const int M = g.Nz / 2;
  const long nrows = (long)g.Nx * g.Ny;
  const long row0 = (long)blockIdx.x * RB;
  const TI* src = in + blockIdx.y * in_arr_stride;
  cplx<T>* dst = out + blockIdx.y * out_arr_stride;
  const int tid = threadIdx.x;
  batched<4, cplx<T>>(
      RB * M,
      [&](int i) {
        const int r = i / M, m = i % M;
        const long row = row0 + r;
        cplx<T> v = {T(0), T(0)};
        if (row < nrows) {
          const TI* p = src + row * g.Nz + 2 * m;
          v = {T(p[0]), T(p[1])};
        }
        return v;
      },
      [&](int i, cplx<T> v) { buf[i] = v; });
  __syncthreads();
  lds_fft<T, MAXPTS, true, false>(buf, radM, RB, 1, M, twM, tid, kThreads);
  const int H = M / 2 + 1;
  for (int i = tid; i < RB * H; i += kThreads) {
    const int r = i / H, k = i % H;
    const long row = row0 + r;
    if (row < nrows) store_row_r2c(buf + r * M, dst + row * g.Zp, M, twN, k);
  }
}

// z pass, spectrum -> real (unnormalised, times `scale`), narr arrays.
template <typename T, typename TO, int MAXPTS>
__global__ __launch_bounds__(kThreads) void k_rows_c2r(const cplx<T>* __restrict__ in, long in_arr_stride,
                                                       TO* __restrict__ out, long out_arr_stride, Geo g, Radices radM,
                                                       const cplx<T>* __restrict__ twM, const cplx<T>* __restrict__ twN,
                                                       int RB, double scale) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  const int M = g.Nz / 2;
  const long nrows = (long)g.Nx * g.Ny;
  const long row0 = (long)blockIdx.x * RB;
  const cplx<T>* src = in + blockIdx.y * in_arr_stride;
  TO* dst = out + blockIdx.y * out_arr_stride;
  const int tid = threadIdx.x;
  const int H = M / 2 + 1;
  batched<4, c2r_item<T>>(
      RB * H,
      [&](int i) {
        const int r = i / H, k = i % H;
        const long row = row0 + r;
        return row < nrows ? fetch_c2r(src + row * g.Zp, M, twN, k) : c2r_item<T>{};
      },
      [&](int i, c2r_item<T> it) {
        const int r = i / H, k = i % H;
        if (row0 + r < nrows) put_c2r(buf + r * M, M, k, it);
        else zero_c2r(buf + r * M, M, k);
      });
  __syncthreads();
  lds_fft<T, MAXPTS, false, false>(buf, radM, RB, 1, M, twM, tid, kThreads);
  for (int i = tid; i < RB * M; i += kThreads) {
    const int r = i / M, m = i % M;
    const long row = row0 + r;
    if (row < nrows) {
      const cplx<T> v = buf[i];
      TO* p = dst + row * g.Nz + 2 * m;
      p[0] = TO(double(v.x) * scale);
      p[1] = TO(double(v.y) * scale);
    }
  }
}

// ---------------------------------------------------------------------------
// In-place column FFTs along x (axis 0) or y (axis 1) on tiles of TW
// consecutive kz.  grid = (ceil(Zh/TW), outer, narr).
// ---------------------------------------------------------------------------
template <typename T, int MAXPTS, bool FWD>
__global__ __launch_bounds__(kThreads) void k_cols(cplx<T>* __restrict__ data, long arr_stride, Geo g, int axis,
                                                   Radices rad, const cplx<T>* __restrict__ tw, int TW) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  const int kz0 = blockIdx.x * TW;
  const int o = blockIdx.y;
  cplx<T>* d = data + blockIdx.z * arr_stride;
  const int L = rad.L;
  const long es = axis == 0 ? (long)g.Ny * g.Zp : g.Zp;
  const long base = axis == 0 ? (long)o * g.Zp : (long)o * g.Ny * g.Zp;
  const int tid = threadIdx.x;
  batched<16, cplx<T>>(
      L * TW,
      [&](int i) {
        const int j = i / TW, kz = kz0 + i % TW;
        return kz < g.Zh ? d[base + j * es + kz] : cplx<T>{T(0), T(0)};
      },
      [&](int i, cplx<T> v) { buf[i] = v; });
  __syncthreads();
  lds_fft<T, MAXPTS, FWD, true>(buf, rad, TW, TW, 1, tw, tid, kThreads);
  for (int i = tid; i < L * TW; i += kThreads) {
    const int j = i / TW, t = i % TW;
    const int kz = kz0 + t;
    if (kz < g.Zh) d[base + j * es + kz] = buf[i];
  }
}

// Multiply by a weight spectrum stored as one real number per k: even weights
// are real, odd (vector) weights purely imaginary (bit a of odd_mask).
template <typename T>
__device__ __forceinline__ cplx<T> wmul(cplx<T> v, T w, bool odd) {
  return odd ? cplx<T>{-v.y * w, v.x * w} : cplx<T>{v.x * w, v.y * w};
}

// x pass of rho_k, then for each of nconv weights: n_a(k) = rho_k w_a(k) and
// its inverse x pass.  grid = (ceil(Zh/TW), Ny).
template <typename T, int MAXPTS>
__global__ __launch_bounds__(kThreads) void k_conv_fwd(const cplx<T>* __restrict__ A, cplx<T>* __restrict__ B,
                                                       long arr_stride, const T* __restrict__ W, int nconv,
                                                       unsigned odd_mask, Geo g, Radices radX,
                                                       const cplx<T>* __restrict__ twX, int TW) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int L = radX.L;
  cplx<T>* bufR = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* bufW = bufR + L * TW;
  const int kz0 = blockIdx.x * TW;
  const int y = blockIdx.y;
  const long es = (long)g.Ny * g.Zp;
  const long base = (long)y * g.Zp;
  const int tid = threadIdx.x;
  batched<8, cplx<T>>(
      L * TW,
      [&](int i) {
        const int j = i / TW, kz = kz0 + i % TW;
        return kz < g.Zh ? A[base + j * es + kz] : cplx<T>{T(0), T(0)};
      },
      [&](int i, cplx<T> v) { bufR[i] = v; });
  __syncthreads();
  lds_fft<T, MAXPTS, true, true>(bufR, radX, TW, TW, 1, twX, tid, kThreads);
  for (int a = 0; a < nconv; ++a) {
    const T* Wa = W + a * arr_stride;
    const bool odd = (odd_mask >> a) & 1u;
    batched<8, T>(
        L * TW,
        [&](int i) {
          const int j = i / TW, kz = kz0 + i % TW;
          return kz < g.Zh ? Wa[base + j * es + kz] : T(0);
        },
        [&](int i, T w) { bufW[i] = wmul(bufR[i], w, odd); });
    __syncthreads();
    lds_fft<T, MAXPTS, false, true>(bufW, radX, TW, TW, 1, twX, tid, kThreads);
    cplx<T>* Ba = B + a * arr_stride;
    for (int i = tid; i < L * TW; i += kThreads) {
      const int j = i / TW, t = i % TW;
      const int kz = kz0 + t;
      if (kz < g.Zh) Ba[base + j * es + kz] = bufW[i];
    }
    __syncthreads();
  }
}

// x pass of each dPhi_a spectrum, acc = sum_a w_a(k) dPhi_a(k), inverse x pass
// of acc.  grid = (ceil(Zh/TW), Ny).
template <typename T, int MAXPTS>
__global__ __launch_bounds__(kThreads) void k_conv_bwd(const cplx<T>* __restrict__ D, cplx<T>* __restrict__ Acc,
                                                       long arr_stride, const T* __restrict__ W, int m,
                                                       unsigned odd_mask, Geo g, Radices radX,
                                                       const cplx<T>* __restrict__ twX, int TW) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int L = radX.L;
  cplx<T>* bufA = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* bufW = bufA + L * TW;
  const int kz0 = blockIdx.x * TW;
  const int y = blockIdx.y;
  const long es = (long)g.Ny * g.Zp;
  const long base = (long)y * g.Zp;
  const int tid = threadIdx.x;
  for (int i = tid; i < L * TW; i += kThreads) bufA[i] = {T(0), T(0)};
  for (int a = 0; a < m; ++a) {
    const cplx<T>* Da = D + a * arr_stride;
    batched<8, cplx<T>>(
        L * TW,
        [&](int i) {
          const int j = i / TW, kz = kz0 + i % TW;
          return kz < g.Zh ? Da[base + j * es + kz] : cplx<T>{T(0), T(0)};
        },
        [&](int i, cplx<T> v) { bufW[i] = v; });
    __syncthreads();
    lds_fft<T, MAXPTS, true, true>(bufW, radX, TW, TW, 1, twX, tid, kThreads);
    const T* Wa = W + a * arr_stride;
    const bool odd = (odd_mask >> a) & 1u;
    batched<8, T>(
        L * TW,
        [&](int i) {
          const int j = i / TW, kz = kz0 + i % TW;
          return kz < g.Zh ? Wa[base + j * es + kz] : T(0);
        },
        [&](int i, T w) {
          const cplx<T> p = wmul(bufW[i], w, odd);
          bufA[i] = {bufA[i].x + p.x, bufA[i].y + p.y};
        });
    __syncthreads();
  }
  lds_fft<T, MAXPTS, false, true>(bufA, radX, TW, TW, 1, twX, tid, kThreads);
  for (int i = tid; i < L * TW; i += kThreads) {
    const int j = i / TW, t = i % TW;
    const int kz = kz0 + t;
    if (kz < g.Zh) Acc[base + j * es + kz] = bufA[i];
  }
}

// ---------------------------------------------------------------------------
// Per row: inverse z of the nconv weighted-density arrays, Phi and dPhi at every
// site, forward z of the m dPhi arrays written back in place.  The mean-field
// array (index m, if present) is written out as a real row (vmf).
// partials: per block {sum Phi}; flag |= 1 if eta > 1 (src/FMT.cpp:289-290).
// ---------------------------------------------------------------------------
template <typename T, int MODEL, int MAXPTS>
__global__ __launch_bounds__(kThreads) void k_z_pointwise(cplx<T>* __restrict__ B, long arr_stride, int m, int nconv,
                                                          double* __restrict__ vmf, Geo g, Radices radM,
                                                          const cplx<T>* __restrict__ twM,
                                                          const cplx<T>* __restrict__ twN, int RB, PhiParams pp,
                                                          double2* __restrict__ partials, int* __restrict__ flag) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  const int M = g.Nz / 2;
  const long nrows = (long)g.Nx * g.Ny;
  const long row0 = (long)blockIdx.x * RB;
  const int tid = threadIdx.x;
  const int H = M / 2 + 1;
  // 1. load + c2r pre-processing for all arrays and rows
  batched<8, c2r_item<T>>(
      nconv * RB * H,
      [&](int i) {
        const int k = i % H, r = (i / H) % RB, a = i / (H * RB);
        const long row = row0 + r;
        return row < nrows ? fetch_c2r(B + a * arr_stride + row * g.Zp, M, twN, k) : c2r_item<T>{};
      },
      [&](int i, c2r_item<T> it) {
        const int k = i % H, r = (i / H) % RB, a = i / (H * RB);
        cplx<T>* lrow = buf + (a * RB + r) * M;
        if (row0 + r < nrows) put_c2r(lrow, M, k, it);
        else zero_c2r(lrow, M, k);
      });
  __syncthreads();
  lds_fft<T, MAXPTS, false, false>(buf, radM, nconv * RB, 1, M, twM, tid, kThreads);
  // 2. pointwise
  CSum F{0.0, 0.0};
  int bad = 0;
  const T* bufT = reinterpret_cast<const T*>(buf);  // real view: row r of array a = bufT[(a*RB+r)*2M + z]
  T* bufTw = reinterpret_cast<T*>(buf);
  const bool has_mf = nconv > m;
  for (int s = tid; s < RB * 2 * M; s += kThreads) {
    const int r = s / (2 * M), z = s % (2 * M);
    const long row = row0 + r;
    if (row >= nrows) continue;
    double n[11];
#pragma unroll
    for (int a = 0; a < 11; ++a) n[a] = a < m ? double(bufT[(long)(a * RB + r) * 2 * M + z]) : 0.0;
    double out[11];
    const double phi = phi_site<MODEL>(n, pp, out);
    if (n[0] > 0.5 && 1 - n[0] < 0.0) bad = 1;
    F.add(phi);
#pragma unroll
    for (int a = 0; a < 11; ++a)
      if (a < m) bufTw[(long)(a * RB + r) * 2 * M + z] = T(out[a]);
    if (has_mf) vmf[row * g.Nz + z] = double(bufT[(long)(m * RB + r) * 2 * M + z]);
  }
  if (bad) atomicOr(flag, 1);
  __syncthreads();
  // 3. forward z of the m dPhi rows, back into B
  lds_fft<T, MAXPTS, true, false>(buf, radM, m * RB, 1, M, twM, tid, kThreads);
  for (int i = tid; i < m * RB * H; i += kThreads) {
    const int k = i % H;
    const int r = (i / H) % RB;
    const int a = i / (H * RB);
    const long row = row0 + r;
    if (row < nrows) store_row_r2c(buf + (a * RB + r) * M, B + a * arr_stride + row * g.Zp, M, twN, k);
  }
  CSum v[1] = {F};
  block_reduce_store<1>(v, partials, blockIdx.x);
}

// Final z pass: dF = [dF_in] + dV ln rho + dV c2r(acc) + v_mf + dV V_ext - dV mu,
// in the order of src/DFT.cpp:377-425.  partials per block:
// {sum (rho ln rho - rho) dV, sum rho v_mf, sum V_ext rho, sum rho}.
// Per-member block of the mixed path's mean split (kMS doubles): [0] sum rho, [1 + a] the mean of weighted density
// a (a < nconv, the mean-field convolution last), [13 + a] dPhi_a at those means (a < m), [25] the FMT force of the
// means, dV N sum_a dPhi_a(nbar) w_a(k = 0).
constexpr int kMS = 32;
constexpr int kMSn = 1, kMSc = 13, kMSk = 25;

struct FinalParams {
  double dV, mu;
  int ideal, fmt, mf, ext, accumulate;
  int hess;  // Hessian-vector mode of k_z_final2 (the V_ext pointer carries v)
  const double* ms;  // mean split (mixed path): the constant [25] of the member joins the FMT force
};

template <typename T, int MAXPTS>
__global__ __launch_bounds__(kThreads) void k_z_final(const cplx<T>* __restrict__ Acc, const double* __restrict__ rho,
                                                      const double* __restrict__ vext, const double* __restrict__ vmf,
                                                      double* __restrict__ dF, Geo g, Radices radM,
                                                      const cplx<T>* __restrict__ twM, const cplx<T>* __restrict__ twN,
                                                      int RB, FinalParams fp, double2* __restrict__ partials) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  const int M = g.Nz / 2;
  const long nrows = (long)g.Nx * g.Ny;
  const long row0 = (long)blockIdx.x * RB;
  const int tid = threadIdx.x;
  const int H = M / 2 + 1;
  if (fp.fmt) {
    batched<4, c2r_item<T>>(
        RB * H,
        [&](int i) {
          const int r = i / H, k = i % H;
          const long row = row0 + r;
          return row < nrows ? fetch_c2r(Acc + row * g.Zp, M, twN, k) : c2r_item<T>{};
        },
        [&](int i, c2r_item<T> it) {
          const int r = i / H, k = i % H;
          if (row0 + r < nrows) put_c2r(buf + r * M, M, k, it);
          else zero_c2r(buf + r * M, M, k);
        });
    __syncthreads();
    lds_fft<T, MAXPTS, false, false>(buf, radM, RB, 1, M, twM, tid, kThreads);
  }
  const T* bufT = reinterpret_cast<const T*>(buf);
  CSum q[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  struct Site {
    double rho, df, vm, ve;
  };
  const long nsite = (long)RB * 2 * M;
  batched<4, Site>(
      (int)nsite,
      [&](int s) {
        const long row = row0 + s / (2 * M);
        Site v{1.0, 0.0, 0.0, 0.0};
        if (row < nrows) {
          const long p = row * g.Nz + s % (2 * M);
          v.rho = rho[p];
          if (fp.accumulate) v.df = dF[p];
          if (fp.mf) v.vm = vmf[p];
          if (fp.ext) v.ve = vext[p];
        }
        return v;
      },
      [&](int s, Site v) {
        const int r = s / (2 * M), z = s % (2 * M);
        const long row = row0 + r;
        if (row >= nrows) return;
        const long p = row * g.Nz + z;
        const double d0 = v.rho;
        double d = v.df;
        if (fp.ideal) {
          const double l = log(d0);
          q[0].add((d0 * l - d0) * fp.dV);
          d += l * fp.dV;
        }
        if (fp.fmt) d += double(bufT[(long)r * 2 * M + z]) * fp.dV;
        if (fp.mf) {
          d += v.vm;
          q[1].add(d0 * v.vm);
        }
        if (fp.ext) {
          q[2].add(v.ve * d0);
          d += v.ve * fp.dV;
        }
        q[3].add(d0);
        d += -fp.mu * fp.dV;
        dF[p] = d;
      });
  block_reduce_store<4>(q, partials, blockIdx.x);
}

// Weight spectra: real part (even weights) or imaginary part (odd weights) of
// r2c(w_a), times `scale`, converted to the pipeline precision.
template <typename T>
static __global__ void k_extract_weight(const cplx<double>* __restrict__ S, T* __restrict__ W, long n, int odd, double scale) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx<double> v = S[i];
  W[i] = T((odd ? v.y : v.x) * scale);
}

static __global__ void k_sum_real(const double* __restrict__ x, long n, double2* __restrict__ partials) {
  CSum a[1] = {{0.0, 0.0}};
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads) a[0].add(x[i]);
  block_reduce_store<1>(a, partials, blockIdx.x);
}

}  // namespace cdft
