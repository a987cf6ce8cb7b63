// Workgroup-level mixed-radix Stockham FFTs held in LDS, for CDNA4 (gfx950).
//
// A workgroup transforms `nb` independent sequences of length L that live in
// LDS, element j of sequence b at buf[b*bs + j*es].  Each pass is one radix-R
// Stockham autosort step (radix 2, 3, 4 or 8): every thread loads up to
// MAXPTS/R butterflies into registers, the workgroup synchronises, then each
// thread applies its twiddles and radix-R DFTs and writes the results back in
// place.  Twiddles W_L^j = exp(-2 pi i j / L) come from a per-length table that
// the host computes in extended precision.
//
// Conventions follow FFTW (the reference's DFT_FFT, include/DFT_LinAlg.h:163-288):
// forward = exp(-2 pi i jk/L), inverse = exp(+2 pi i jk/L), both unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace cdft {

template <typename T>
struct cplx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return {a.x, -a.y}; }
// multiply by -i*s  (s = +1 forward, -1 inverse): (x, y) -> (s*y, -s*x)
template <typename T, int S>
__device__ __forceinline__ cplx<T> mul_mi(cplx<T> a) { return S > 0 ? cplx<T>{a.y, -a.x} : cplx<T>{-a.y, a.x}; }

// Length descriptor: radix sequence of one 1D transform (host-built).
struct Radices {
  int L;
  int npass;
  int r[12];
};

// -------------------------------------------------------------------------
// Small DFTs in registers.  FWD = true uses exp(-2 pi i /R).
// -------------------------------------------------------------------------
template <typename T, bool FWD>
__device__ __forceinline__ void dft2(cplx<T>* v) {
  cplx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T, bool FWD>
__device__ __forceinline__ void dft4(cplx<T>* v) {
  constexpr int S = FWD ? 1 : -1;
  cplx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  cplx<T> a2 = cadd(v[1], v[3]), a3 = mul_mi<T, S>(csub(v[1], v[3]));
  v[0] = cadd(a0, a2);
  v[2] = csub(a0, a2);
  v[1] = cadd(a1, a3);
  v[3] = csub(a1, a3);
}

template <typename T, bool FWD>
__device__ __forceinline__ void dft3(cplx<T>* v) {
  // y1,2 = v0 - (v1+v2)/2 -/+ i*s*(sqrt3/2)(v1-v2), s = +1 forward
  const T h = T(0.86602540378443864676372317075294);
  cplx<T> s = cadd(v[1], v[2]);
  cplx<T> d = csub(v[1], v[2]);
  cplx<T> t = {v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y};
  // -i*d*h (forward) = (h*d.y, -h*d.x)
  cplx<T> r = FWD ? cplx<T>{h * d.y, -h * d.x} : cplx<T>{-h * d.y, h * d.x};
  v[0] = cadd(v[0], s);
  v[1] = cadd(t, r);
  v[2] = csub(t, r);
}

template <typename T, bool FWD>
__device__ __forceinline__ void dft8(cplx<T>* v) {
  constexpr int S = FWD ? 1 : -1;
  const T c = T(0.70710678118654752440084436210485);
  // stage 1: radix-2 across (k, k+4)
  cplx<T> a[8];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = cadd(v[k], v[k + 4]);
    a[k + 4] = csub(v[k], v[k + 4]);
  }
  // twiddles w8^k on the difference half: 1, (1 -/+ i)/sqrt2, -/+ i, (-1 -/+ i)/sqrt2
  {
    cplx<T> t = a[5];
    a[5] = FWD ? cplx<T>{c * (t.x + t.y), c * (t.y - t.x)} : cplx<T>{c * (t.x - t.y), c * (t.y + t.x)};
    a[6] = mul_mi<T, S>(a[6]);
    t = a[7];
    a[7] = FWD ? cplx<T>{c * (t.y - t.x), -c * (t.x + t.y)} : cplx<T>{-c * (t.x + t.y), c * (t.x - t.y)};
  }
  // two radix-4 DFTs on (a0..a3) and (a4..a7); outputs interleave even/odd
  dft4<T, FWD>(a);
  dft4<T, FWD>(a + 4);
  v[0] = a[0]; v[2] = a[1]; v[4] = a[2]; v[6] = a[3];
  v[1] = a[4]; v[3] = a[5]; v[5] = a[6]; v[7] = a[7];
}

template <typename T, int R, bool FWD>
__device__ __forceinline__ void dftR(cplx<T>* v) {
  if constexpr (R == 2) dft2<T, FWD>(v);
  else if constexpr (R == 3) dft3<T, FWD>(v);
  else if constexpr (R == 4) dft4<T, FWD>(v);
  else dft8<T, FWD>(v);
}

template <typename T, bool FWD>
__device__ __forceinline__ cplx<T> twiddle(const cplx<T>* __restrict__ tw, int idx) {
  cplx<T> w = tw[idx];
  return FWD ? w : cconj(w);
}

// One Stockham pass (Bainville's indexing): for butterfly i of a sequence,
// k = i mod p, inputs x[i + r*m] (m = L/R) are twiddled by W_L^{r*k*L/(p*R)},
// transformed, and written to y[(i-k)*R + k + r*p].
// BFAST: consecutive threads take consecutive sequences (column tiles) rather
// than consecutive butterflies of one sequence (rows).
template <typename T, int R, int MAXPTS, bool FWD, bool BFAST>
__device__ __forceinline__ void stockham_pass(cplx<T>* buf, int L, int nb, int es, int bs, int p,
                                              const cplx<T>* __restrict__ tw, int tid, int nthr) {
  constexpr int MB = (MAXPTS + R - 1) / R;
  const int m = L / R;
  const int total = m * nb;
  const int tstep = L / (p * R);
  cplx<T> v[MB][R];
  int dst[MB];
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int idx = tid + q * nthr;
    dst[q] = -1;
    if (idx < total) {
      int b, i;
      if (BFAST) { b = idx % nb; i = idx / nb; }
      else { i = idx % m; b = idx / m; }
      const int k = i % p;
      cplx<T>* base = buf + b * bs;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = base[(i + r * m) * es];
      if (k) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[q][r] = cmul(v[q][r], twiddle<T, FWD>(tw, r * k * tstep));
      }
      dst[q] = b * bs + ((i - k) * R + k) * es;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    if (dst[q] >= 0) {
      dftR<T, R, FWD>(v[q]);
      cplx<T>* o = buf + dst[q];
      const int pes = p * es;
#pragma unroll
      for (int r = 0; r < R; ++r) o[r * pes] = v[q][r];
    }
  }
  __syncthreads();
}

// Full transform of nb sequences of length rad.L in LDS.  Requires
// (L/R)*nb*R <= nthr*MAXPTS for every radix R used, i.e. L*nb <= nthr*MAXPTS.
template <typename T, int MAXPTS, bool FWD, bool BFAST>
__device__ void lds_fft(cplx<T>* buf, const Radices& rad, int nb, int es, int bs, const cplx<T>* __restrict__ tw,
                        int tid, int nthr) {
  int p = 1;
  for (int s = 0; s < rad.npass; ++s) {
    const int R = rad.r[s];
    if (R == 8) stockham_pass<T, 8, MAXPTS, FWD, BFAST>(buf, rad.L, nb, es, bs, p, tw, tid, nthr);
    else if (R == 4) stockham_pass<T, 4, MAXPTS, FWD, BFAST>(buf, rad.L, nb, es, bs, p, tw, tid, nthr);
    else if (R == 2) stockham_pass<T, 2, MAXPTS, FWD, BFAST>(buf, rad.L, nb, es, bs, p, tw, tid, nthr);
    else stockham_pass<T, 3, MAXPTS, FWD, BFAST>(buf, rad.L, nb, es, bs, p, tw, tid, nthr);
    p *= R;
  }
}

}  // namespace cdft
