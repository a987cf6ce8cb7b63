// Learned warm start: a 3-level 3D U-Net on periodic grids, (beta V_ext, mu) ->
// log rho, with hand-written BF16 MFMA kernels (FP32 accumulation) for the
// forward pass and both backward products, trained on the device with Adam.
//
// Layout: activations NDHWC in BF16, T[b][x][y][z][c] with c fastest (C a
// multiple of 8, so 8 channels are one 16-byte load); weights [co][tap][ci]
// in BF16 with tap = 9 (dx+1) + 3 (dy+1) + (dz+1), K = 27 * CIN; FP32 master
// weights, gradients and Adam moments.  Convolutions are 3x3x3 with periodic
// wrap, like the systems the network is trained on.
//
// MFMA: v_mfma_f32_16x16x32_bf16, D(16x16) += A(16x32) B(32x16); lane l holds
// A[l%16][8(l/16)+j], B[8(l/16)+j][l%16] and D[4(l/16)+r][l%16].
//   forward / backward-data:  rows = 16 consecutive z positions, cols = 16
//     output channels, K = (tap, input channel);
//   weight gradient:          rows = 16 output channels, cols = 16 input
//     channels of one tap, K = positions (per-chunk partial sums, reduced in a
//     fixed order).
#include <cstdint>
#include <random>
#include <vector>

#include "system.hpp"

namespace cdft_impl {
namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int wrap(int i, int D) { return i < 0 ? i + D : (i >= D ? i - D : i); }

// Y[p][co] = act(bias[co] + sum_{tap,ci} W[co][tap][ci] X[p + off(tap)][ci]).  One wavefront per MT tiles of 16
// consecutive positions, all COUT channels; blocks of 4 wavefronts.  Each weight fragment a wavefront loads feeds
// the MT tiles' MFMAs and each tap is decoded once for them; every output keeps the same K order (the result does
// not depend on MT).
template <int CIN, int COUT, bool RELU, int MT>
__global__ __launch_bounds__(256) void k_conv3(const bf16* __restrict__ X, const bf16* __restrict__ W,
                                               const float* __restrict__ bias, bf16* __restrict__ Y, int D, long ntiles) {
  static_assert(CIN % 8 == 0 && COUT % 16 == 0, "channel counts");
  constexpr int K = 27 * CIN, NCB = COUT / 16;
  const long tile0 = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * MT;
  if (tile0 >= ntiles) return;
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
  const long P = (long)D * D * D;
  // row r of tile mt: position (tile0 + mt) 16 + r (a tile may span rows and samples on coarse grids)
  const bf16* Xb[MT];
  int xs[MT], ys[MT], zs[MT];
  bool live[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    live[mt] = tile0 + mt < ntiles;
    const long p = (live[mt] ? tile0 + mt : tile0) * 16 + r;
    const long b = p / P;
    const int q = (int)(p - b * P);
    xs[mt] = q / (D * D);
    ys[mt] = (q / D) % D;
    zs[mt] = q % D;
    Xb[mt] = X + b * P * CIN;
  }
  f32x4 acc[MT][NCB];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) acc[mt][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    const int k = k0 + 8 * g;
    const bool kin = k < K;
    const int tap = kin ? k / CIN : 0, c = k - tap * CIN;
    const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
    bf16x8 w[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      if (kin) {
        w[cb] = *reinterpret_cast<const bf16x8*>(W + (long)(cb * 16 + r) * K + k);
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) w[cb][j] = (bf16)0.f;
      }
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      bf16x8 a;
      if (kin) {
        const long src =
            ((long)wrap(xs[mt] + dx, D) * D + wrap(ys[mt] + dy, D)) * D + wrap(zs[mt] + dz, D);
        a = *reinterpret_cast<const bf16x8*>(Xb[mt] + src * CIN + c);
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = (bf16)0.f;
      }
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) acc[mt][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w[cb], acc[mt][cb], 0, 0, 0);
    }
  }
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    if (!live[mt]) continue;
    const long p0 = (tile0 + mt) * 16;
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      const int co = cb * 16 + r;
      const float bv = bias ? bias[co] : 0.f;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        float v = acc[mt][cb][rr] + bv;
        if (RELU) v = fmaxf(v, 0.f);
        Y[(p0 + 4 * g + rr) * COUT + co] = (bf16)v;
      }
    }
  }
}

// Tiles per wavefront of k_conv3: as many (4, 2, 1) as keep at least 1024 blocks of 4 wavefronts in the launch.
template <typename F>
void with_conv_mt(long ntiles, F&& f) {
  if (ntiles >= 4L * 4 * 1024) f(std::integral_constant<int, 4>{});
  else if (ntiles >= 2L * 4 * 1024) f(std::integral_constant<int, 2>{});
  else f(std::integral_constant<int, 1>{});
}

// Partial weight gradients: part[chunk][co][tap][ci] = sum over the chunk's positions of
// dZ[p][co] X[p + off(tap)][ci].  Wavefront job = (16 co, tap, 16 ci); chunk = `chunk` positions.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad3(const bf16* __restrict__ X, const bf16* __restrict__ dZ,
                                                float* __restrict__ part, int D, long npos, int chunk) {
  constexpr int NCI = (CIN + 15) / 16, NJOB = (COUT / 16) * 27 * NCI;
  const int job = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (job >= NJOB) return;
  const int cb = job / (27 * NCI), rem = job % (27 * NCI), tap = rem / NCI, nb = rem % NCI;
  const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
  const int co = cb * 16 + r, ci = nb * 16 + r;
  const long P = (long)D * D * D;
  const long pbeg = (long)blockIdx.x * chunk;
  const long pend = pbeg + chunk < npos ? pbeg + chunk : npos;
  f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long p0 = pbeg; p0 < pend; p0 += 32) {
    const long pg = p0 + 8 * g;  // 8 consecutive z of one row (chunks and every level's grid are multiples of 8)
    bf16x8 a, bb;
    if (pg < pend) {
      const long b = pg / P;
      const int q = (int)(pg - b * P);
      const int x = q / (D * D), y = (q / D) % D, z = q % D;
      const long row = ((long)wrap(x + dx, D) * D + wrap(y + dy, D)) * D;
      const bf16* Xr = X + b * P * CIN;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        a[j] = dZ[(pg + j) * COUT + co];
        bb[j] = ci < CIN ? Xr[(row + wrap(z + j + dz, D)) * CIN + ci] : (bf16)0.f;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] = bb[j] = (bf16)0.f;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc, 0, 0, 0);
  }
  // D[row = co block row 4g+rr][col = ci r]
  float* pc = part + (long)blockIdx.x * COUT * 27 * CIN;
  if (ci < CIN) {
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) pc[((long)(cb * 16 + 4 * g + rr) * 27 + tap) * CIN + ci] = acc[rr];
  }
}

// grad[i] = sum_c part[c][i] (fixed order); bias gradient from bpart[c][co].
__global__ void k_wgrad_reduce(const float* __restrict__ part, int nchunk, long n, float* __restrict__ grad) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(long)c * n + i];
    grad[i] = s;
  }
}

// Bias gradient partials: bpart[blk][co] = sum over the block's positions of dZ[p][co].
template <int COUT>
__global__ void k_bgrad(const bf16* __restrict__ dZ, long npos, int chunk, float* __restrict__ bpart) {
  const int co = threadIdx.x % COUT, lane_set = threadIdx.x / COUT, nset = blockDim.x / COUT;
  const long pbeg = (long)blockIdx.x * chunk, pend = pbeg + chunk < npos ? pbeg + chunk : npos;
  float s = 0.f;
  for (long p = pbeg + lane_set; p < pend; p += nset) s += (float)dZ[p * COUT + co];
  __shared__ float sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (lane_set == 0) {
    float t = 0.f;
    for (int k = 0; k < nset; ++k) t += sh[k * COUT + co];
    bpart[(long)blockIdx.x * COUT + co] = t;
  }
}

// dZ = (dA [+ dB]) * [Y > 0] elementwise (ReLU backward, optionally summing two gradient sources).
// (dA and dZ may be the same buffer)
__global__ void k_relu_bwd(const bf16* dA, const bf16* __restrict__ dB, const bf16* __restrict__ Y, bf16* dZ, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float v = (float)dA[i];
    if (dB) v += (float)dB[i];
    dZ[i] = (float)Y[i] > 0.f ? (bf16)v : (bf16)0.f;
  }
}

// 2x2x2 average pooling, D -> D/2, C channels.
__global__ void k_pool2(const bf16* __restrict__ X, bf16* __restrict__ Y, int D, int C, long n_out) {
  const int h = D / 2;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n_out; i += (long)gridDim.x * blockDim.x) {
    const int c = i % C;
    long q = i / C;
    const int z = q % h, y = (q / h) % h, x = (q / ((long)h * h)) % h;
    const long b = q / ((long)h * h * h);
    float s = 0.f;
    for (int d = 0; d < 8; ++d)
      s += (float)X[((((b * D + 2 * x + (d >> 2)) * D) + 2 * y + ((d >> 1) & 1)) * D + 2 * z + (d & 1)) * C + c];
    Y[i] = (bf16)(0.125f * s);
  }
}

// Pooling backward: dX (D) = dY (D/2) / 8 at every child.
__global__ void k_pool2_bwd(const bf16* __restrict__ dY, bf16* __restrict__ dX, int D, int C, long n_in) {
  const int h = D / 2;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n_in; i += (long)gridDim.x * blockDim.x) {
    const int c = i % C;
    long q = i / C;
    const int z = q % D, y = (q / D) % D, x = (q / ((long)D * D)) % D;
    const long b = q / ((long)D * D * D);
    dX[i] = (bf16)(0.125f * (float)dY[((((b * h + x / 2) * h) + y / 2) * h + z / 2) * C + c]);
  }
}

// Nearest upsampling D/2 -> D of C channels into channels [off, off + C) of a Ct-channel tensor.
__global__ void k_up2(const bf16* __restrict__ X, bf16* __restrict__ Y, int D, int C, int Ct, int off, long n_out) {
  const int h = D / 2;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n_out; i += (long)gridDim.x * blockDim.x) {
    const int c = i % C;
    long q = i / C;
    const int z = q % D, y = (q / D) % D, x = (q / ((long)D * D)) % D;
    const long b = q / ((long)D * D * D);
    Y[q * Ct + off + c] = X[((((b * h + x / 2) * h) + y / 2) * h + z / 2) * C + c];
  }
}

// Upsampling backward: dX (D/2, C) = sum over the 8 children of dY channels [off, off + C) (Ct channels).
__global__ void k_up2_bwd(const bf16* __restrict__ dY, bf16* __restrict__ dX, int D, int C, int Ct, int off, long n_in) {
  const int h = D / 2;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n_in; i += (long)gridDim.x * blockDim.x) {
    const int c = i % C;
    long q = i / C;
    const int z = q % h, y = (q / h) % h, x = (q / ((long)h * h)) % h;
    const long b = q / ((long)h * h * h);
    float s = 0.f;
    for (int d = 0; d < 8; ++d)
      s += (float)dY[((((b * D + 2 * x + (d >> 2)) * D) + 2 * y + ((d >> 1) & 1)) * D + 2 * z + (d & 1)) * Ct + off + c];
    dX[i] = (bf16)s;
  }
}

// Channel block copy between a C-channel tensor a and channels [off, off + C) of a
// Ct-channel tensor b: a -> b (a_to_b = 1) or b -> a.
__global__ void k_chan_copy(bf16* __restrict__ a, bf16* __restrict__ b, int C, int Ct, int off, long n, int a_to_b) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long p = i / C;
    const int c = i % C;
    if (a_to_b) b[p * Ct + off + c] = a[i];
    else a[i] = b[p * Ct + off + c];
  }
}

// Network input: channel 0 = beta V_ext / 3, channel 1 = (mu + 3) / 2, channels 2..7 = 0.
__global__ void k_input(const float* __restrict__ V, const int* __restrict__ idx, const float* __restrict__ mu,
                        bf16* __restrict__ X, long P, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long b = i / P, p = i - b * P;
    const int s = idx[b];
    bf16* o = X + i * 8;
    o[0] = (bf16)(V[(long)s * P + p] * (1.f / 3.f));
    o[1] = (bf16)((mu[s] + 3.f) * 0.5f);
#pragma unroll
    for (int c = 2; c < 8; ++c) o[c] = (bf16)0.f;
  }
}

// Head: out[p] = hb + sum_c hw[c] Y[p][c] (FP32).
__global__ void k_head(const bf16* __restrict__ Y, const float* __restrict__ hw, const float* __restrict__ hb,
                       float* __restrict__ out, long n) {
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    float s = hb[0];
#pragma unroll
    for (int c = 0; c < 16; ++c) s += hw[c] * (float)Y[p * 16 + c];
    out[p] = s;
  }
}

// Loss and head backward: d = 2 (out - t) / n; dY[p][c] = hw[c] d; partial sums of d^2 n/4 (loss),
// d Y[p][c] (head weights) and d (head bias) per block.
__global__ void k_loss_head_bwd(const float* __restrict__ out, const float* __restrict__ T, const int* __restrict__ idx,
                                long P, const bf16* __restrict__ Y, const float* __restrict__ hw,
                                bf16* __restrict__ dY, float* __restrict__ part, long n) {
  float acc[18];
#pragma unroll
  for (int k = 0; k < 18; ++k) acc[k] = 0.f;
  const float inv = 1.f / (float)n;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    const long b = p / P, q = p - b * P;
    const float e = out[p] - T[(long)idx[b] * P + q];
    const float d = 2.f * e * inv;
    acc[16] += e * e;
    acc[17] += d;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const float yc = (float)Y[p * 16 + c];
      acc[c] += d * yc;
      dY[p * 16 + c] = (bf16)(hw[c] * d);
    }
  }
  __shared__ float sh[18][256];
#pragma unroll
  for (int k = 0; k < 18; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  if (threadIdx.x < 18) {
    float s = 0.f;
    for (int t = 0; t < (int)blockDim.x; ++t) s += sh[threadIdx.x][t];
    part[blockIdx.x * 18 + threadIdx.x] = s;
  }
}

__global__ void k_sum_parts(const float* __restrict__ part, int nblk, int nq, float* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= nq) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += part[b * nq + k];
  out[k] = s;
}

// Adam on the FP32 master parameters (bias-corrected), then the BF16 copies.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m, float* __restrict__ v,
                       long n, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = gr[i];
    const float mi = b1 * m[i] + (1 - b1) * gi, vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// BF16 forward weights [co][tap][ci] and backward-data weights [ci][26 - tap][co] from the FP32 master.
__global__ void k_weights_bf16(const float* __restrict__ w, bf16* __restrict__ wf, bf16* __restrict__ wt, int cin,
                               int cout, int cin_stored) {
  const long n = (long)cout * 27 * cin;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int ci = i % cin, tap = (i / cin) % 27, co = i / (27L * cin);
    const bf16 v = (bf16)w[i];
    wf[((long)co * 27 + tap) * cin_stored + ci] = v;
    if (wt) wt[((long)ci * 27 + (26 - tap)) * cout + co] = v;
  }
}

}  // namespace
}  // namespace cdft_impl

using namespace cdft_impl;

namespace {

struct LayerSpec {
  int cin, cout, level;  // level 0: D, 1: D/2, 2: D/4
};
// enc1a enc1b | enc2a enc2b | bottleneck a b | dec2a dec2b | dec1a dec1b
const LayerSpec kLayers[10] = {{8, 16, 0},  {16, 16, 0}, {16, 32, 1}, {32, 32, 1}, {32, 64, 2},
                               {64, 64, 2}, {96, 32, 1}, {32, 32, 1}, {48, 16, 0}, {16, 16, 0}};
constexpr int kWChunk = 4096;  // positions per weight-gradient partial

template <typename F>
void dispatch_layer(int li, F&& f) {
  switch (li) {
    case 0: f(std::integral_constant<int, 8>{}, std::integral_constant<int, 16>{}); break;
    case 1: case 9: f(std::integral_constant<int, 16>{}, std::integral_constant<int, 16>{}); break;
    case 2: f(std::integral_constant<int, 16>{}, std::integral_constant<int, 32>{}); break;
    case 3: case 7: f(std::integral_constant<int, 32>{}, std::integral_constant<int, 32>{}); break;
    case 4: f(std::integral_constant<int, 32>{}, std::integral_constant<int, 64>{}); break;
    case 5: f(std::integral_constant<int, 64>{}, std::integral_constant<int, 64>{}); break;
    case 6: f(std::integral_constant<int, 96>{}, std::integral_constant<int, 32>{}); break;
    default: f(std::integral_constant<int, 48>{}, std::integral_constant<int, 16>{}); break;
  }
}

inline int gs(long n) { return (int)std::min<long>(4096, (n + 255) / 256); }

}  // namespace

struct cdft_unet_s {
  int D = 64, B = 0;  // grid, batch capacity
  hipStream_t stream = nullptr;
  long nparam = 0;
  long woff[10]{}, boff[10]{}, hoff = 0;  // offsets in the flat FP32 parameter vector
  DevBuf w, g, m, v;                      // FP32 master, gradient, Adam moments
  DevBuf wf[10], wt[10];                  // BF16 forward / backward-data weights
  // activations (BF16) and their gradients, sized for B samples
  DevBuf x0, y[10], p1, p2, u1, u2, out, dy[10], du1, du2, dp1, dp2, dskip1, dskip3, dpool1, dpool3;
  DevBuf idx, wpart, bpart, lpart, lsum;
  int step = 0;

  ~cdft_unet_s() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  int Dl(int level) const { return D >> level; }
  long npos(int level) const { return (long)B * Dl(level) * Dl(level) * Dl(level); }

  void init_params(unsigned long long seed) {
    long o = 0;
    for (int i = 0; i < 10; ++i) {
      woff[i] = o;
      o += (long)kLayers[i].cout * 27 * kLayers[i].cin;
      boff[i] = o;
      o += kLayers[i].cout;
    }
    hoff = o;
    o += 17;
    nparam = o;
    std::vector<float> h(nparam, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int i = 0; i < 10; ++i) {
      const double sd = std::sqrt(2.0 / (27.0 * (i == 0 ? 2 : kLayers[i].cin)));  // He; layer 0 sees 2 channels
      for (long k = 0; k < (long)kLayers[i].cout * 27 * kLayers[i].cin; ++k) h[woff[i] + k] = (float)(sd * nd(rng));
    }
    for (int c = 0; c < 16; ++c) h[hoff + c] = (float)(0.25 * nd(rng));
    w.alloc(sizeof(float) * nparam);
    g.alloc(sizeof(float) * nparam);
    m.alloc(sizeof(float) * nparam);
    v.alloc(sizeof(float) * nparam);
    HIPCHK(hipMemcpy(w.p, h.data(), sizeof(float) * nparam, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(m.p, 0, sizeof(float) * nparam));
    HIPCHK(hipMemset(v.p, 0, sizeof(float) * nparam));
    for (int i = 0; i < 10; ++i) {
      const long n = (long)kLayers[i].cout * 27 * kLayers[i].cin;
      wf[i].alloc(sizeof(bf16) * n);
      wt[i].alloc(sizeof(bf16) * n);
    }
    refresh_bf16();
  }

  void refresh_bf16() {
    for (int i = 0; i < 10; ++i) {
      const long n = (long)kLayers[i].cout * 27 * kLayers[i].cin;
      hipLaunchKernelGGL(k_weights_bf16, dim3(gs(n)), dim3(256), 0, stream, w.as<float>() + woff[i], wf[i].as<bf16>(),
                         wt[i].as<bf16>(), kLayers[i].cin, kLayers[i].cout, kLayers[i].cin);
    }
    HIPCHK(hipGetLastError());
  }

  void ensure_batch(int nb) {
    if (nb == B) return;
    B = nb;
    x0.alloc(sizeof(bf16) * npos(0) * 8);
    for (int i = 0; i < 10; ++i) {
      y[i].alloc(sizeof(bf16) * npos(kLayers[i].level) * kLayers[i].cout);
      dy[i].alloc(sizeof(bf16) * npos(kLayers[i].level) * kLayers[i].cout);
    }
    p1.alloc(sizeof(bf16) * npos(1) * 16);
    p2.alloc(sizeof(bf16) * npos(2) * 32);
    u2.alloc(sizeof(bf16) * npos(1) * 96);
    u1.alloc(sizeof(bf16) * npos(0) * 48);
    du2.alloc(sizeof(bf16) * npos(1) * 96);
    du1.alloc(sizeof(bf16) * npos(0) * 48);
    dp1.alloc(sizeof(bf16) * npos(1) * 16);
    dp2.alloc(sizeof(bf16) * npos(2) * 32);
    dskip1.alloc(sizeof(bf16) * npos(0) * 16);
    dskip3.alloc(sizeof(bf16) * npos(1) * 32);
    dpool1.alloc(sizeof(bf16) * npos(0) * 16);
    dpool3.alloc(sizeof(bf16) * npos(1) * 32);
    out.alloc(sizeof(float) * npos(0));
    idx.alloc(sizeof(int) * B);
    long maxw = 0;
    for (int i = 0; i < 10; ++i)
      maxw = std::max(maxw, (npos(kLayers[i].level) + kWChunk - 1) / kWChunk * kLayers[i].cout * 27 * kLayers[i].cin);
    wpart.alloc(sizeof(float) * maxw);
    bpart.alloc(sizeof(float) * ((npos(0) + kWChunk - 1) / kWChunk) * 64);
    lpart.alloc(sizeof(float) * 1024 * 18);
    lsum.alloc(sizeof(float) * 18);
  }

  void conv(int li, const bf16* X, bf16* Y, bool relu) {
    const int D_ = Dl(kLayers[li].level);
    const long ntiles = npos(kLayers[li].level) / 16;
    dispatch_layer(li, [&](auto ci, auto co) {
      constexpr int CI = decltype(ci)::value, CO = decltype(co)::value;
      with_conv_mt(ntiles, [&](auto mtc) {
        constexpr int MT = decltype(mtc)::value;
        const dim3 grid((ntiles + 4 * MT - 1) / (4 * MT));
        if (relu)
          hipLaunchKernelGGL((k_conv3<CI, CO, true, MT>), grid, dim3(256), 0, stream, X, wf[li].as<bf16>(),
                             w.as<float>() + boff[li], Y, D_, ntiles);
        else
          hipLaunchKernelGGL((k_conv3<CI, CO, false, MT>), grid, dim3(256), 0, stream, X, wf[li].as<bf16>(),
                             w.as<float>() + boff[li], Y, D_, ntiles);
      });
    });
    HIPCHK(hipGetLastError());
  }

  // dX = conv^T(dZ) (layers 1..9)
  void conv_bwd_data(int li, const bf16* dZ, bf16* dX) {
    const int D_ = Dl(kLayers[li].level);
    const long ntiles = npos(kLayers[li].level) / 16;
    dispatch_layer(li, [&](auto ci, auto co) {
      constexpr int CI = decltype(ci)::value, CO = decltype(co)::value;
      if constexpr (CI % 16 == 0)
        with_conv_mt(ntiles, [&](auto mtc) {
          constexpr int MT = decltype(mtc)::value;
          hipLaunchKernelGGL((k_conv3<CO, CI, false, MT>), dim3((ntiles + 4 * MT - 1) / (4 * MT)), dim3(256), 0,
                             stream, dZ, wt[li].as<bf16>(), nullptr, dX, D_, ntiles);
        });
    });
    HIPCHK(hipGetLastError());
  }

  // weight and bias gradients of layer li into g
  void conv_wgrad(int li, const bf16* X, const bf16* dZ) {
    const int D_ = Dl(kLayers[li].level);
    const long np = npos(kLayers[li].level);
    const int nchunk = (int)((np + kWChunk - 1) / kWChunk);
    dispatch_layer(li, [&](auto ci, auto co) {
      constexpr int CI = decltype(ci)::value, CO = decltype(co)::value;
      constexpr int njob = (CO / 16) * 27 * ((CI + 15) / 16);
      hipLaunchKernelGGL((k_wgrad3<CI, CO>), dim3(nchunk, (njob + 3) / 4), dim3(256), 0, stream, X, dZ,
                         wpart.as<float>(), D_, np, kWChunk);
      const long n = (long)CO * 27 * CI;
      hipLaunchKernelGGL(k_wgrad_reduce, dim3(gs(n)), dim3(256), 0, stream, wpart.as<float>(), nchunk, n,
                         g.as<float>() + woff[li]);
      hipLaunchKernelGGL((k_bgrad<CO>), dim3(nchunk), dim3(256 / CO * CO), 0, stream, dZ, np, kWChunk,
                         bpart.as<float>());
      hipLaunchKernelGGL(k_wgrad_reduce, dim3(1), dim3(64), 0, stream, bpart.as<float>(), nchunk, (long)CO,
                         g.as<float>() + boff[li]);
    });
    HIPCHK(hipGetLastError());
  }

  void relu_bwd(const bf16* dA, const bf16* dB, const bf16* Y, bf16* dZ, long n) {
    hipLaunchKernelGGL(k_relu_bwd, dim3(gs(n)), dim3(256), 0, stream, dA, dB, Y, dZ, n);
  }

  // forward from the prepared input x0 -> out (FP32 log rho)
  void forward() {
    const long n0 = npos(0), n1 = npos(1), n2 = npos(2);
    conv(0, x0.as<bf16>(), y[0].as<bf16>(), true);
    conv(1, y[0].as<bf16>(), y[1].as<bf16>(), true);
    hipLaunchKernelGGL(k_pool2, dim3(gs(n1 * 16)), dim3(256), 0, stream, y[1].as<bf16>(), p1.as<bf16>(), D, 16, n1 * 16);
    conv(2, p1.as<bf16>(), y[2].as<bf16>(), true);
    conv(3, y[2].as<bf16>(), y[3].as<bf16>(), true);
    hipLaunchKernelGGL(k_pool2, dim3(gs(n2 * 32)), dim3(256), 0, stream, y[3].as<bf16>(), p2.as<bf16>(), D / 2, 32,
                       n2 * 32);
    conv(4, p2.as<bf16>(), y[4].as<bf16>(), true);
    conv(5, y[4].as<bf16>(), y[5].as<bf16>(), true);
    hipLaunchKernelGGL(k_up2, dim3(gs(n1 * 64)), dim3(256), 0, stream, y[5].as<bf16>(), u2.as<bf16>(), D / 2, 64, 96, 0,
                       n1 * 64);
    hipLaunchKernelGGL(k_chan_copy, dim3(gs(n1 * 32)), dim3(256), 0, stream, y[3].as<bf16>(), u2.as<bf16>(), 32, 96, 64,
                       n1 * 32, 1);
    conv(6, u2.as<bf16>(), y[6].as<bf16>(), true);
    conv(7, y[6].as<bf16>(), y[7].as<bf16>(), true);
    hipLaunchKernelGGL(k_up2, dim3(gs(n0 * 32)), dim3(256), 0, stream, y[7].as<bf16>(), u1.as<bf16>(), D, 32, 48, 0,
                       n0 * 32);
    hipLaunchKernelGGL(k_chan_copy, dim3(gs(n0 * 16)), dim3(256), 0, stream, y[1].as<bf16>(), u1.as<bf16>(), 16, 48, 32,
                       n0 * 16, 1);
    conv(8, u1.as<bf16>(), y[8].as<bf16>(), true);
    conv(9, y[8].as<bf16>(), y[9].as<bf16>(), true);
    hipLaunchKernelGGL(k_head, dim3(gs(n0)), dim3(256), 0, stream, y[9].as<bf16>(), w.as<float>() + hoff,
                       w.as<float>() + hoff + 16, out.as<float>(), n0);
    HIPCHK(hipGetLastError());
  }

  // loss, backward, gradients in g; returns nothing (loss in lsum[16] / n)
  void backward(const float* T) {
    const long n0 = npos(0), n1 = npos(1), n2 = npos(2);
    const int nblk = 1024;
    const long P = (long)D * D * D;
    hipLaunchKernelGGL(k_loss_head_bwd, dim3(nblk), dim3(256), 0, stream, out.as<float>(), T, idx.as<int>(), P,
                       y[9].as<bf16>(), w.as<float>() + hoff, dy[9].as<bf16>(), lpart.as<float>(), n0);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(32), 0, stream, lpart.as<float>(), nblk, 18, lsum.as<float>());
    // head gradients: lsum[0..15] = d hw, lsum[17] = d hb
    HIPCHK(hipMemcpyAsync(g.as<float>() + hoff, lsum.p, sizeof(float) * 16, hipMemcpyDeviceToDevice, stream));
    HIPCHK(hipMemcpyAsync(g.as<float>() + hoff + 16, lsum.as<float>() + 17, sizeof(float), hipMemcpyDeviceToDevice,
                          stream));
    bf16* dz = dy[9].as<bf16>();  // reuse: dZ9 in place
    relu_bwd(dy[9].as<bf16>(), nullptr, y[9].as<bf16>(), dz, n0 * 16);
    conv_wgrad(9, y[8].as<bf16>(), dz);
    conv_bwd_data(9, dz, dy[8].as<bf16>());
    relu_bwd(dy[8].as<bf16>(), nullptr, y[8].as<bf16>(), dy[8].as<bf16>(), n0 * 16);
    conv_wgrad(8, u1.as<bf16>(), dy[8].as<bf16>());
    conv_bwd_data(8, dy[8].as<bf16>(), du1.as<bf16>());
    // split du1: [0,32) -> up(y7), [32,48) -> skip y1
    hipLaunchKernelGGL(k_up2_bwd, dim3(gs(n1 * 32)), dim3(256), 0, stream, du1.as<bf16>(), dy[7].as<bf16>(), D, 32, 48, 0,
                       n1 * 32);
    hipLaunchKernelGGL(k_chan_copy, dim3(gs(n0 * 16)), dim3(256), 0, stream, dskip1.as<bf16>(), du1.as<bf16>(), 16, 48,
                       32, n0 * 16, 0);
    relu_bwd(dy[7].as<bf16>(), nullptr, y[7].as<bf16>(), dy[7].as<bf16>(), n1 * 32);
    conv_wgrad(7, y[6].as<bf16>(), dy[7].as<bf16>());
    conv_bwd_data(7, dy[7].as<bf16>(), dy[6].as<bf16>());
    relu_bwd(dy[6].as<bf16>(), nullptr, y[6].as<bf16>(), dy[6].as<bf16>(), n1 * 32);
    conv_wgrad(6, u2.as<bf16>(), dy[6].as<bf16>());
    conv_bwd_data(6, dy[6].as<bf16>(), du2.as<bf16>());
    hipLaunchKernelGGL(k_up2_bwd, dim3(gs(n2 * 64)), dim3(256), 0, stream, du2.as<bf16>(), dy[5].as<bf16>(), D / 2, 64,
                       96, 0, n2 * 64);
    hipLaunchKernelGGL(k_chan_copy, dim3(gs(n1 * 32)), dim3(256), 0, stream, dskip3.as<bf16>(), du2.as<bf16>(), 32, 96,
                       64, n1 * 32, 0);
    relu_bwd(dy[5].as<bf16>(), nullptr, y[5].as<bf16>(), dy[5].as<bf16>(), n2 * 64);
    conv_wgrad(5, y[4].as<bf16>(), dy[5].as<bf16>());
    conv_bwd_data(5, dy[5].as<bf16>(), dy[4].as<bf16>());
    relu_bwd(dy[4].as<bf16>(), nullptr, y[4].as<bf16>(), dy[4].as<bf16>(), n2 * 64);
    conv_wgrad(4, p2.as<bf16>(), dy[4].as<bf16>());
    conv_bwd_data(4, dy[4].as<bf16>(), dp2.as<bf16>());
    hipLaunchKernelGGL(k_pool2_bwd, dim3(gs(n1 * 32)), dim3(256), 0, stream, dp2.as<bf16>(), dpool3.as<bf16>(), D / 2,
                       32, n1 * 32);
    relu_bwd(dpool3.as<bf16>(), dskip3.as<bf16>(), y[3].as<bf16>(), dy[3].as<bf16>(), n1 * 32);
    conv_wgrad(3, y[2].as<bf16>(), dy[3].as<bf16>());
    conv_bwd_data(3, dy[3].as<bf16>(), dy[2].as<bf16>());
    relu_bwd(dy[2].as<bf16>(), nullptr, y[2].as<bf16>(), dy[2].as<bf16>(), n1 * 32);
    conv_wgrad(2, p1.as<bf16>(), dy[2].as<bf16>());
    conv_bwd_data(2, dy[2].as<bf16>(), dp1.as<bf16>());
    hipLaunchKernelGGL(k_pool2_bwd, dim3(gs(n0 * 16)), dim3(256), 0, stream, dp1.as<bf16>(), dpool1.as<bf16>(), D, 16,
                       n0 * 16);
    relu_bwd(dpool1.as<bf16>(), dskip1.as<bf16>(), y[1].as<bf16>(), dy[1].as<bf16>(), n0 * 16);
    conv_wgrad(1, y[0].as<bf16>(), dy[1].as<bf16>());
    conv_bwd_data(1, dy[1].as<bf16>(), dy[0].as<bf16>());
    relu_bwd(dy[0].as<bf16>(), nullptr, y[0].as<bf16>(), dy[0].as<bf16>(), n0 * 16);
    conv_wgrad(0, x0.as<bf16>(), dy[0].as<bf16>());
    HIPCHK(hipGetLastError());
  }

  void prepare(const float* V, const float* mu, const int* idx_host, int nb) {
    ensure_batch(nb);
    HIPCHK(hipMemcpyAsync(idx.p, idx_host, sizeof(int) * nb, hipMemcpyHostToDevice, stream));
    const long P = (long)D * D * D, n = P * nb;
    hipLaunchKernelGGL(k_input, dim3(gs(n)), dim3(256), 0, stream, V, idx.as<int>(), mu, x0.as<bf16>(), P, n);
    HIPCHK(hipGetLastError());
  }
};

extern "C" {

int cdft_unet_create(int D, unsigned long long seed, cdft_unet* out) {
  return guarded([&] {
    REQUIRE(out && D >= 32 && D % 32 == 0, "grid must be a multiple of 32");
    auto u = std::make_unique<cdft_unet_s>();
    u->D = D;
    HIPCHK(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking));
    u->init_params(seed);
    HIPCHK(hipStreamSynchronize(u->stream));
    *out = u.release();
    return (int)CDFT_OK;
  });
}

void cdft_unet_destroy(cdft_unet u) { delete u; }

int cdft_unet_num_params(cdft_unet u, long long* n) {
  return guarded([&] {
    REQUIRE(u && n, "null argument");
    *n = u->nparam;
    return (int)CDFT_OK;
  });
}

int cdft_unet_get_params(cdft_unet u, float* host) {
  return guarded([&] {
    REQUIRE(u && host, "null argument");
    HIPCHK(hipMemcpyAsync(host, u->w.p, sizeof(float) * u->nparam, hipMemcpyDeviceToHost, u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    return (int)CDFT_OK;
  });
}

int cdft_unet_set_params(cdft_unet u, const float* host) {
  return guarded([&] {
    REQUIRE(u && host, "null argument");
    HIPCHK(hipMemcpyAsync(u->w.p, host, sizeof(float) * u->nparam, hipMemcpyHostToDevice, u->stream));
    u->refresh_bf16();
    HIPCHK(hipStreamSynchronize(u->stream));
    return (int)CDFT_OK;
  });
}

int cdft_unet_forward(cdft_unet u, const float* V_dev, const float* mu_dev, const int* idx_host, int nb, float* out_dev) {
  return guarded([&] {
    REQUIRE(u && V_dev && mu_dev && idx_host && out_dev && nb > 0, "bad argument");
    u->prepare(V_dev, mu_dev, idx_host, nb);
    u->forward();
    HIPCHK(hipMemcpyAsync(out_dev, u->out.p, sizeof(float) * u->npos(0), hipMemcpyDeviceToDevice, u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    return (int)CDFT_OK;
  });
}

int cdft_unet_grad(cdft_unet u, const float* V_dev, const float* mu_dev, const float* T_dev, const int* idx_host, int nb,
                   float* loss, float* grad_host) {
  return guarded([&] {
    REQUIRE(u && V_dev && mu_dev && T_dev && idx_host && nb > 0, "bad argument");
    u->prepare(V_dev, mu_dev, idx_host, nb);
    u->forward();
    u->backward(T_dev);
    float ls[18];
    HIPCHK(hipMemcpyAsync(ls, u->lsum.p, sizeof(ls), hipMemcpyDeviceToHost, u->stream));
    if (grad_host) HIPCHK(hipMemcpyAsync(grad_host, u->g.p, sizeof(float) * u->nparam, hipMemcpyDeviceToHost, u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    if (loss) *loss = ls[16] / (float)u->npos(0);
    return (int)CDFT_OK;
  });
}

int cdft_unet_train_step(cdft_unet u, const float* V_dev, const float* mu_dev, const float* T_dev, const int* idx_host,
                         int nb, float lr, float* loss) {
  return guarded([&] {
    REQUIRE(u && V_dev && mu_dev && T_dev && idx_host && nb > 0, "bad argument");
    u->prepare(V_dev, mu_dev, idx_host, nb);
    u->forward();
    u->backward(T_dev);
    u->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)u->step), c2 = 1.f - std::pow(b2, (float)u->step);
    hipLaunchKernelGGL(k_adam, dim3(gs(u->nparam)), dim3(256), 0, u->stream, u->w.as<float>(), u->g.as<float>(),
                       u->m.as<float>(), u->v.as<float>(), u->nparam, lr, b1, b2, 1e-8f, c1, c2);
    u->refresh_bf16();
    if (loss) {
      float ls[18];
      HIPCHK(hipMemcpyAsync(ls, u->lsum.p, sizeof(ls), hipMemcpyDeviceToHost, u->stream));
      HIPCHK(hipStreamSynchronize(u->stream));
      *loss = ls[16] / (float)u->npos(0);
    }
    return (int)CDFT_OK;
  });
}

}  // extern "C"
