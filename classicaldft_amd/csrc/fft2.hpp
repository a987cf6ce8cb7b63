// Register-resident two-pass FFTs (four-step, L = R1 * R2) for CDNA4.
//
// With n = R2 n1 + n2 and k = k1 + R1 k2:
//   X[k1 + R1 k2] = sum_n2 W_R2^{n2 k2} W_L^{n2 k1} sum_n1 x[R2 n1 + n2] W_R1^{n1 k1}.
// P = gcd(R1, R2) threads cooperate on one sequence.  Pass 1: each thread takes
// Q1 = R2/P values of n2, holds the R1 inputs x[R2 n1 + n2] in registers (loaded
// straight from HBM by the caller), runs a radix-R1 DFT and the twiddle
// W_L^{n2 k1}; one LDS exchange; pass 2: each thread takes Q2 = R1/P values of
// k1 and runs a radix-R2 DFT over n2, leaving X[k1 + R1 k2] in registers.
// An inverse transform can start directly from that register layout by using
// the swapped factorisation (R2, R1): the held index k1 plays the role of n2.
//
// Compared with a per-radix Stockham sweep through LDS this needs one LDS round
// trip and one barrier per transform, and every thread issues R1 independent
// global loads at once.  Small DFTs are unrolled with constant twiddles.
#pragma once
#include <hip/hip_runtime.h>

#include "fft.hpp"

namespace cdft {

// cos / sin (2 pi j / 48), j = 0..47, covers every radix used here.
__device__ constexpr double kC48[48] = {
    1.0, 0.99144486137381041114, 0.96592582628906828675, 0.92387953251128675613, 0.86602540378443864676,
    0.79335334029123516458, 0.70710678118654752440, 0.60876142900872063942, 0.5, 0.38268343236508977173,
    0.25881904510252076235, 0.13052619222005159155, 0.0, -0.13052619222005159155, -0.25881904510252076235,
    -0.38268343236508977173, -0.5, -0.60876142900872063942, -0.70710678118654752440, -0.79335334029123516458,
    -0.86602540378443864676, -0.92387953251128675613, -0.96592582628906828675, -0.99144486137381041114, -1.0,
    -0.99144486137381041114, -0.96592582628906828675, -0.92387953251128675613, -0.86602540378443864676,
    -0.79335334029123516458, -0.70710678118654752440, -0.60876142900872063942, -0.5, -0.38268343236508977173,
    -0.25881904510252076235, -0.13052619222005159155, 0.0, 0.13052619222005159155, 0.25881904510252076235,
    0.38268343236508977173, 0.5, 0.60876142900872063942, 0.70710678118654752440, 0.79335334029123516458,
    0.86602540378443864676, 0.92387953251128675613, 0.96592582628906828675, 0.99144486137381041114};

// W_N^j = exp(-/+ 2 pi i j / N) for compile-time N | 48 and constant j.
template <typename T, int N, bool FWD>
__device__ __forceinline__ cplx<T> wconst(int j) {
  const int a = (j % N) * (48 / N);
  const T c = T(kC48[a]);
  const T s = T(kC48[(a + 36) % 48]);  // sin(x) = cos(x - pi/2)
  return FWD ? cplx<T>{c, -s} : cplx<T>{c, s};
}

template <typename T, int R, bool FWD>
__device__ __forceinline__ void dft_small(cplx<T>* v);

// Generic small DFT by four-step recursion on constant factors.
template <typename T, int A, int B, bool FWD>
__device__ __forceinline__ void dft_split(cplx<T>* v) {
  // L = A * B, n = B n1 + n2, k = k1 + A k2
  constexpr int L = A * B;
  cplx<T> t[B][A];
#pragma unroll
  for (int n2 = 0; n2 < B; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) t[n2][n1] = v[B * n1 + n2];
    dft_small<T, A, FWD>(t[n2]);
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1)
      if (n2) t[n2][k1] = cmul(t[n2][k1], wconst<T, L, FWD>(n2 * k1));
  }
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    cplx<T> u[B];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) u[n2] = t[n2][k1];
    dft_small<T, B, FWD>(u);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = u[k2];
  }
}

template <typename T, int R, bool FWD>
__device__ __forceinline__ void dft_small(cplx<T>* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) dft2<T, FWD>(v);
  else if constexpr (R == 3) dft3<T, FWD>(v);
  else if constexpr (R == 4) dft4<T, FWD>(v);
  else if constexpr (R == 8) dft8<T, FWD>(v);
  else if constexpr (R == 6) dft_split<T, 2, 3, FWD>(v);
  else if constexpr (R == 12) dft_split<T, 4, 3, FWD>(v);
  else if constexpr (R == 16) dft_split<T, 4, 4, FWD>(v);
  else static_assert(R == 16, "unsupported radix");
}

constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }

// Factorisations L = R1 * R2 used by the register FFT (f2_r1 also serves the host's twiddle tables).
constexpr int f2_r1(int L) {
  return L == 256 ? 16 : L == 128 ? 16 : L == 96 ? 12 : L == 64 ? 8 : L == 48 ? 12 : L == 32 ? 8 : L == 24 ? 8
         : L == 16 ? 4 : L == 12 ? 4 : L == 8 ? 4 : L == 6 ? 3 : L == 4 ? 2 : L;
}

template <int L>
struct F2 {
  static constexpr int R1 = f2_r1(L);
  static constexpr int R2 = L / R1;
  static constexpr int P = cgcd(R1, R2);
  static constexpr int Q1 = R2 / P;  // pass-1 sequences (values of n2) per thread
  static constexpr int Q2 = R1 / P;  // pass-2 sequences (values of k1) per thread
  static_assert(R1 * R2 == L, "unsupported length");
};

template <int L>
constexpr bool reg_fft_supported() {
  return L == 256 || L == 128 || L == 96 || L == 64 || L == 48 || L == 32 || L == 24 || L == 16 || L == 12 ||
         L == 8 || L == 6 || L == 4;
}

// One transform of a sequence shared by P threads (lane t in [0, P)).
// Input  v1[q][n1] = x[R2 n1 + n2], n2 = t + P q   (q < Q1)
// Output v2[q][k2] = X[k1 + R1 k2], k1 = t + P q   (q < Q2)
// lds: this sequence's exchange slot, element (k1, n2) at lds[(k1 (R2 + PAD) + n2) ls]
// (PAD = 1 breaks the power-of-two stride of the pass-2 reads when ls = 1);
// tw: W_L^j table (j < L).  The caller must __syncthreads() before reusing lds.
// The two halves of reg_fft_seq; the caller synchronises the block between them.
template <typename T, int L, int R1, int R2, bool FWD, int PAD = 0>
__device__ __forceinline__ void reg_fft_pass1(cplx<T> (&v1)[R2 / cgcd(R1, R2)][R1], int t, cplx<T>* lds, int ls,
                                              const cplx<T>* __restrict__ tw) {
  constexpr int P = cgcd(R1, R2), Q1 = R2 / P, S = R2 + PAD;
#pragma unroll
  for (int q = 0; q < Q1; ++q) {
    const int n2 = t + P * q;
    dft_small<T, R1, FWD>(v1[q]);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      cplx<T> x = v1[q][k1];
      if (k1) x = cmul(x, twiddle<T, FWD>(tw, n2 * k1));
      lds[(k1 * S + n2) * ls] = x;
    }
  }
}

// reg_fft_pass1 with its twiddles in pass order: twp[(q R1 + k1) P + t] = W_L^{(t + P q) k1} (the host's
// upload_pass_twiddles), so the P lanes of a sequence read consecutive entries at constant offsets (no LDS bank
// conflicts when the table sits in LDS) instead of the strided W_L^{n2 k1}.  Same values, same results.
template <typename T, int L, int R1, int R2, bool FWD, int PAD = 0>
__device__ __forceinline__ void reg_fft_pass1p(cplx<T> (&v1)[R2 / cgcd(R1, R2)][R1], int t, cplx<T>* lds, int ls,
                                               const cplx<T>* __restrict__ twp) {
  constexpr int P = cgcd(R1, R2), Q1 = R2 / P, S = R2 + PAD;
  const cplx<T>* tt = twp + t;
#pragma unroll
  for (int q = 0; q < Q1; ++q) {
    const int n2 = t + P * q;
    dft_small<T, R1, FWD>(v1[q]);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      cplx<T> x = v1[q][k1];
      if (k1) {
        const cplx<T> w = tt[(q * R1 + k1) * P];
        x = cmul(x, FWD ? w : cconj(w));
      }
      lds[(k1 * S + n2) * ls] = x;
    }
  }
}

template <typename T, int L, int R1, int R2, bool FWD, int PAD = 0>
__device__ __forceinline__ void reg_fft_pass2(cplx<T> (&v2)[R1 / cgcd(R1, R2)][R2], int t, const cplx<T>* lds,
                                              int ls) {
  constexpr int P = cgcd(R1, R2), Q2 = R1 / P, S = R2 + PAD;
#pragma unroll
  for (int q = 0; q < Q2; ++q) {
    const int k1 = t + P * q;
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) v2[q][n2] = lds[(k1 * S + n2) * ls];
    dft_small<T, R2, FWD>(v2[q]);
  }
}

template <typename T, int L, int R1, int R2, bool FWD, int PAD = 0>
__device__ __forceinline__ void reg_fft_seq(cplx<T> (&v1)[R2 / cgcd(R1, R2)][R1],
                                            cplx<T> (&v2)[R1 / cgcd(R1, R2)][R2], int t, cplx<T>* lds, int ls,
                                            const cplx<T>* __restrict__ tw) {
  reg_fft_pass1<T, L, R1, R2, FWD, PAD>(v1, t, lds, ls, tw);
  __syncthreads();
  reg_fft_pass2<T, L, R1, R2, FWD, PAD>(v2, t, lds, ls);
}

// reg_fft_seq with the pass-ordered pass-1 table of reg_fft_pass1p.
template <typename T, int L, int R1, int R2, bool FWD, int PAD = 0>
__device__ __forceinline__ void reg_fft_seqp(cplx<T> (&v1)[R2 / cgcd(R1, R2)][R1],
                                             cplx<T> (&v2)[R1 / cgcd(R1, R2)][R2], int t, cplx<T>* lds, int ls,
                                             const cplx<T>* __restrict__ twp) {
  reg_fft_pass1p<T, L, R1, R2, FWD, PAD>(v1, t, lds, ls, twp);
  __syncthreads();
  reg_fft_pass2<T, L, R1, R2, FWD, PAD>(v2, t, lds, ls);
}

}  // namespace cdft
