// Internal state shared by the translation units of libcdft: device buffers,
// twiddle tables, the system object and the launch helpers of the transforms.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <cstdlib>
#include <type_traits>

#include "../../include/cdft.h"
#include "kernels.hpp"
#include "kernels2.hpp"
#include "rows2.hpp"
#include "weights.hpp"
#include "weights_ref.hpp"
#include "constraints.hpp"
#include "dst.hpp"

using namespace cdft;

namespace cdft_impl {

extern thread_local std::string g_err;

struct CdftError : std::runtime_error {
  int code;
  CdftError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                                      \
  do {                                                                                                    \
    hipError_t e_ = (expr);                                                                               \
    if (e_ != hipSuccess)                                                                                 \
      throw CdftError(CDFT_HIP_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_) + " (line " +    \
                                          std::to_string(__LINE__) + ")");                                \
  } while (0)

#define REQUIRE(cond, msg) \
  do {                     \
    if (!(cond)) throw CdftError(CDFT_BAD_ARG, msg); \
  } while (0)

// A stream created for one standalone call, destroyed on every exit path.
struct OwnedStream {
  hipStream_t s = nullptr;
  OwnedStream() { HIPCHK(hipStreamCreate(&s)); }
  OwnedStream(const OwnedStream&) = delete;
  OwnedStream& operator=(const OwnedStream&) = delete;
  ~OwnedStream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t b) {
    release();
    if (b) HIPCHK(hipMalloc(&p, b));
    bytes = b;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename U>
  U* as() const { return reinterpret_cast<U*>(p); }
};

inline bool smooth23(int n) {
  if (n < 1) return false;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  return n == 1;
}

inline Radices make_radices(int L) {
  Radices r{};
  r.L = L;
  int m = L;
  int k = 0;
  while (m % 8 == 0) { r.r[k++] = 8; m /= 8; }
  if (m % 4 == 0) { r.r[k++] = 4; m /= 4; }
  if (m % 2 == 0) { r.r[k++] = 2; m /= 2; }
  while (m % 3 == 0) { r.r[k++] = 3; m /= 3; }
  if (m != 1) throw CdftError(CDFT_BAD_ARG, "transform length " + std::to_string(L) + " is not a product of 2 and 3");
  r.npass = k;
  return r;
}

// W_L^j = exp(-2 pi i j / L), j in [0, L), in extended precision.
template <typename T>
inline void upload_twiddles(DevBuf& buf, int L) {
  std::vector<cplx<T>> h(L);
  const long double tp = 6.283185307179586476925286766559005768L;
  for (int j = 0; j < L; ++j) {
    const long double a = -tp * (long double)j / (long double)L;
    h[j] = {T(cosl(a)), T(sinl(a))};
  }
  buf.alloc(sizeof(cplx<T>) * L);
  HIPCHK(hipMemcpy(buf.p, h.data(), sizeof(cplx<T>) * L, hipMemcpyHostToDevice));
}

template <typename Fn>
inline bool with_reg_len(int L, Fn&& f);

// The pass-1 twiddles of the register FFT of length L in pass order (reg_fft_pass1p): entry (q R1 + k1) P + t is
// W_L^{(t + P q) k1} with L = R1 R2, P = gcd(R1, R2), q < R2 / P; the same values as the plain table's entries.
template <typename T>
inline void upload_pass_twiddles(DevBuf& buf, int L) {
  const int R1 = f2_r1(L), R2 = L / R1, P = cgcd(R1, R2), Q1 = R2 / P;
  std::vector<cplx<T>> h(L);
  const long double tp = 6.283185307179586476925286766559005768L;
  for (int q = 0; q < Q1; ++q)
    for (int k1 = 0; k1 < R1; ++k1)
      for (int t = 0; t < P; ++t) {
        const long double a = -tp * (long double)((t + P * q) * k1) / (long double)L;
        h[(q * R1 + k1) * P + t] = {T(cosl(a)), T(sinl(a))};
      }
  buf.alloc(sizeof(cplx<T>) * L);
  HIPCHK(hipMemcpy(buf.p, h.data(), sizeof(cplx<T>) * L, hipMemcpyHostToDevice));
}

template <typename T>
struct Tables {
  DevBuf x, y, m, n;  // Nx, Ny, Nz/2, Nz
  DevBuf mp;          // Nz/2 in pass order (register z rows only)
  void build(const Geo& g) {
    upload_twiddles<T>(x, g.Nx);
    upload_twiddles<T>(y, g.Ny);
    upload_twiddles<T>(m, std::max(1, g.Nz / 2));
    upload_twiddles<T>(n, g.Nz);
    if (g.Nz % 2 == 0 && with_reg_len(g.Nz / 2, [](auto) {})) upload_pass_twiddles<T>(mp, g.Nz / 2);
  }
};

inline int div_up(long a, long b) { return int((a + b - 1) / b); }

// alias <-> density: FMT_Species (c = (1 - 1e-18 * 4 pi d^3/3)/dV) or plain
// Species (rho = 1e-18 + x^2, src/Species.cpp:24-48).
struct Alias {
  double c;
  int fmt;
  __device__ __forceinline__ double rho(double y) const {
    return fmt ? 1e-18 + c * y * y / (1 + y * y) : 1e-18 + y * y;
  }
  __device__ __forceinline__ double x_of(double r) const {
    if (fmt) {
      const double z = (r - 1e-18) / c;
      return sqrt(fabs(z / (1 - z)));
    }
    return sqrt(fmax(0.0, r - 1e-20));
  }
  __device__ __forceinline__ double dfac(double y) const {
    return fmt ? 2 * c * y / ((1 + y * y) * (1 + y * y)) : 2 * y;
  }
};

}  // namespace cdft_impl

using namespace cdft_impl;

struct cdft_system_s {
  cdft_system_desc d{};
  int device = 0;
  hipStream_t stream = nullptr;      // where the system's work runs: own_stream, or one the caller set
  hipStream_t own_stream = nullptr;  // created with the system (null for a mixture's later members)
  Geo g{};
  long N = 0, Nkp = 0, nrows = 0;
  int m = 0, nconv = 0;
  // tensor models whose weights satisfy the trace identity w_Txx + w_Tyy + w_Tzz = w_s (checked at set-up): T_zz is
  // not stored, m = 10, and the pointwise kernels rebuild n_Tzz = n_s - n_Txx - n_Tyy
  int trace = 0;
  unsigned odd_mask = 0x1Cu;  // arrays 2, 3, 4 (vector weights) have imaginary spectra
  bool mixed = false;
  Radices radX{}, radY{}, radM{};
  Tables<double> twd;
  Tables<float> twf;
  DevBuf A, B, W, vmf, rho, dF, vext, partials, partials2, flag, results;
  DevBuf A2, B2, vdir;  // Hessian-vector product: spectra of the direction, its host copy (allocated on first use)
  DevBuf Wao, psi;      // Asakura-Oosawa: polymer weight spectra (m arrays), psi / exp(-psi) (N per member)
  bool wq = false;      // W holds the folded (kx, ky <= N/2) spectra (ColGeo::wq)
  unsigned wq_stride = 0;
  bool has_ao = false;
  double a_vdw = 0.0;
  int RBrow = 1, RBpw = 1, TWcol = 16, TWconv = 8;
  // batch of nb same-shape members (weights shared; rho, dF, V_ext, v_mf, mu,
  // spectra, partial sums, flags and results per member); nb = 1 for one system
  int nb = 1;
  long pslots = 0;  // partial-sum slots per member
  DevBuf mus;       // per-member mu (batches only)
  bool reg_fft = true;  // register two-pass FFT for the x / y passes where the length allows
  // half the columns per block in the register column kernels: x/y passes,
  // forward convolution, backward convolution (tuned per precision at create)
  bool nb_half_cols = false, nb_half_fwd = false, nb_half_bwd = false;
  int pw_rows = 0;       // rows per block of the register pointwise kernel (0: automatic)
  bool pw3 = true;       // specialised pointwise kernel (pw3.hpp) where an instance covers the system
  bool trace_ok = true;  // the trace-reduced weight set may be used (create_system's allow_trace)
  Skip skip;             // stopped members of a FIRE batch, set while its steps are enqueued (kernels2.hpp)
  // mixed path: the means of the weighted densities and of dPhi are carried in FP64 and the FP32 arrays hold the
  // deviations from them (msplit: kMS doubles per member; w0: the FP64 weight spectra at k = 0, nconv values)
  bool mean_split = false;
  DevBuf msplit, w0;
  long slab_bytes = 0;   // cache budget of one x-slab of the y / z / y section (0: no slabs)
  bool has_vext = false;
  bool last_with_dft = true;
  // species constraints (desc fixed_mass / boundary); no_ideal: the next evaluation skips the ideal gas
  bool no_ideal = false, last_ideal = true, last_fixed_mass = false;
  std::vector<double> lagrange;  // fixed-mass multiplier of each member after the last fetched evaluation
  std::vector<double> masses;    // fixed mass of each member (desc fixed_mass unless set per member)
  DevBuf masses_d;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev;
  std::vector<const char*> ev_names;
  int n_marks = 0;

  ~cdft_system_s() {
    for (auto e : ev) (void)hipEventDestroy(e);
    if (own_stream) (void)hipStreamDestroy(own_stream);
  }

  void mark(const char* name) {
    if (!timing) return;
    if (n_marks < (int)ev.size()) {
      HIPCHK(hipEventRecord(ev[n_marks], stream));
      ev_names[n_marks] = name;
      ++n_marks;
    }
  }
};

namespace cdft_impl {

template <typename T>
const Tables<T>& tables(const cdft_system_s& s);
template <>
inline const Tables<double>& tables<double>(const cdft_system_s& s) { return s.twd; }
template <>
inline const Tables<float>& tables<float>(const cdft_system_s& s) { return s.twf; }

// Calls f(std::integral_constant<int, L>{}) when the register FFT handles
// length L; returns false otherwise.
template <typename Fn>
inline bool with_reg_len(int L, Fn&& f) {
  switch (L) {
#define CDFT_REG_LEN(n)                       \
  case n:                                     \
    f(std::integral_constant<int, n>{});      \
    return true;
    CDFT_REG_LEN(256) CDFT_REG_LEN(128) CDFT_REG_LEN(96) CDFT_REG_LEN(64) CDFT_REG_LEN(48) CDFT_REG_LEN(32)
    CDFT_REG_LEN(24) CDFT_REG_LEN(16) CDFT_REG_LEN(12) CDFT_REG_LEN(8) CDFT_REG_LEN(6) CDFT_REG_LEN(4)
#undef CDFT_REG_LEN
    default: return false;
  }
}

// Rows per block of the register z kernels: RB * P <= 256 threads and
// RB * SLOT elements of LDS within cap bytes (0 if one row does not fit).
template <typename T, int M>
constexpr int row_rb(int cap) {
  constexpr int per = RowF<T, M>::SLOT * int(sizeof(cplx<T>)), P = RowF<T, M>::P;
  int rb = kThreads / P;
  const int wave = 64 / P;  // rows per full wavefront
  while (rb > wave && (rb * per > cap || rb % wave)) --rb;
  while (rb > 0 && rb * per > cap) --rb;
  return rb;
}

// The register z kernels address arrays through buffer resources (< 2^31 bytes).
template <typename T>
inline bool rows_reg_ok(const cdft_system_s& s) {
  return s.reg_fft && s.Nkp * (long)sizeof(cplx<T>) < (1L << 31) && s.N * 8L < (1L << 31) &&
         with_reg_len(s.g.Nz / 2, [](auto) {});
}

// ------------------------------- transforms --------------------------------
// psum (register kernels only): per-block FP64 sums of each input array at psum + array * psp (returns the block
// count through nblk_out).
template <typename T, typename TI>
inline void launch_rows_r2c(const cdft_system_s& s, const TI* in, long in_stride, cplx<T>* out, long out_stride, int narr,
                            double2* psum = nullptr, long psp = 0, long* nblk_out = nullptr) {
  const Tables<T>& tw = tables<T>(s);
  if (rows_reg_ok<T>(s)) {
    with_reg_len(s.g.Nz / 2, [&](auto mc) {
      constexpr int Mc = decltype(mc)::value;
      constexpr int RB = row_rb<T, Mc>(32768), P = RowF<T, Mc>::P, SLOT = RowF<T, Mc>::SLOT;
      if (nblk_out) *nblk_out = div_up(s.nrows, RB);
      hipLaunchKernelGGL((k_rows_r2c2<T, TI, Mc>), dim3(div_up(s.nrows, RB), narr), dim3(RB * P),
                         RB * SLOT * sizeof(cplx<T>), s.stream, in, in_stride, out, out_stride, s.nrows, s.g.Zp,
                         tw.mp.template as<cplx<T>>(), tw.n.template as<cplx<T>>(), RB, psum, psp, s.skip);
    });
    HIPCHK(hipGetLastError());
    return;
  }
  REQUIRE(!psum, "row sums need the register z kernels");
  const int M = s.g.Nz / 2;
  const int RB = s.RBrow;
  dim3 grid(div_up(s.nrows, RB), narr);
  hipLaunchKernelGGL((k_rows_r2c<T, TI, 4>), grid, dim3(kThreads), sizeof(cplx<T>) * RB * M, s.stream, in, in_stride,
                     out, out_stride, s.g, s.radM, tw.m.template as<cplx<T>>(), tw.n.template as<cplx<T>>(), RB);
  HIPCHK(hipGetLastError());
}

template <typename T, typename TO>
inline void launch_rows_c2r(const cdft_system_s& s, const cplx<T>* in, long in_stride, TO* out, long out_stride, int narr,
                     double scale) {
  const Tables<T>& tw = tables<T>(s);
  if (rows_reg_ok<T>(s)) {
    with_reg_len(s.g.Nz / 2, [&](auto mc) {
      constexpr int Mc = decltype(mc)::value;
      constexpr int RB = row_rb<T, Mc>(32768), P = RowF<T, Mc>::P, SLOT = RowF<T, Mc>::SLOT;
      hipLaunchKernelGGL((k_rows_c2r2<T, TO, Mc>), dim3(div_up(s.nrows, RB), narr), dim3(RB * P),
                         RB * SLOT * sizeof(cplx<T>), s.stream, in, in_stride, out, out_stride, s.nrows, s.g.Zp,
                         tw.mp.template as<cplx<T>>(), tw.n.template as<cplx<T>>(), RB, scale);
    });
    HIPCHK(hipGetLastError());
    return;
  }
  const int M = s.g.Nz / 2;
  const int RB = s.RBrow;
  dim3 grid(div_up(s.nrows, RB), narr);
  hipLaunchKernelGGL((k_rows_c2r<T, TO, 4>), grid, dim3(kThreads), sizeof(cplx<T>) * RB * M, s.stream, in, in_stride,
                     out, out_stride, s.g, s.radM, tw.m.template as<cplx<T>>(), tw.n.template as<cplx<T>>(), RB,
                     scale);
  HIPCHK(hipGetLastError());
}

// Columns per block of the register FFT kernels: NB * P <= 256 threads and an
// exchange buffer of L * NB elements of at most 64 KiB; `half` halves NB.
template <typename T, int L>
constexpr int reg_nb_max() {
  int nb = kThreads / F2<L>::P;
  while (nb > 1 && nb * L * int(sizeof(cplx<T>)) > 65536) nb /= 2;
  return nb;
}

template <typename T, int L, typename Fn>
inline void with_nb(bool half, Fn&& f) {
  constexpr int nb = reg_nb_max<T, L>();
  if (half && nb > 1)
    f(std::integral_constant<int, nb / 2>{});
  else
    f(std::integral_constant<int, nb>{});
}

// Flat column numbering of the x (axis 0) or y (axis 1) pass.
template <typename T>
inline ColGeo col_geo(const Geo& g, int axis) {
  ColGeo c{};
  c.bytes = unsigned((long)g.Nx * g.Ny * g.Zp * (long)sizeof(cplx<T>));
  if (axis == 0) {
    c.ncols = unsigned(g.Ny * g.Zp);
    c.cz = c.ncols;
    c.os = 0;
    c.es = unsigned(g.Ny * g.Zp);
  } else {
    c.ncols = unsigned(g.Nx * g.Zp);
    c.cz = unsigned(g.Zp);
    c.os = unsigned(g.Ny * g.Zp);
    c.es = unsigned(g.Zp);
  }
  return c;
}

// The register kernels address one spectral array through a buffer resource
// (byte offsets below 2^31).
template <typename T>
inline bool reg_fft_ok(const cdft_system_s& s, int L) {
  return s.reg_fft && s.Nkp * (long)sizeof(cplx<T>) < (1L << 31) && with_reg_len(L, [](auto) {});
}

template <typename T>
inline bool launch_cols_reg(const cdft_system_s& s, cplx<T>* data, long stride, int narr, int axis, bool fwd,
                            int members, long mstride) {
  const int L = axis == 0 ? s.g.Nx : s.g.Ny;
  if (!reg_fft_ok<T>(s, L)) return false;
  const Tables<T>& tw = tables<T>(s);
  const cplx<T>* t = (axis == 0 ? tw.x : tw.y).template as<cplx<T>>();
  ColGeo cg = col_geo<T>(s.g, axis);
  cg.sk = s.skip;
  return with_reg_len(L, [&](auto lc) {
    constexpr int Lc = decltype(lc)::value;
    with_nb<T, Lc>(s.nb_half_cols, [&](auto nbc) {
      constexpr int NB = decltype(nbc)::value;
      const dim3 grid(div_up(cg.ncols, NB), narr, members), block(NB * F2<Lc>::P);
      if (fwd)
        hipLaunchKernelGGL((k_cols2<T, Lc, NB, true>), grid, block, 0, s.stream, data, stride, cg, t, mstride);
      else
        hipLaunchKernelGGL((k_cols2<T, Lc, NB, false>), grid, block, 0, s.stream, data, stride, cg, t, mstride);
      HIPCHK(hipGetLastError());
    });
  });
}

// y pass over the x-planes [x0, x1) only (register kernels): the columns of those planes are one contiguous range
// of the flat column numbering, so the pass is the full-array pass on an offset base with fewer columns.
template <typename T>
inline void launch_cols_slab(const cdft_system_s& s, cplx<T>* data, long stride, int narr, bool fwd, int members,
                             long mstride, int x0, int x1) {
  const Tables<T>& tw = tables<T>(s);
  const cplx<T>* t = tw.y.template as<cplx<T>>();
  ColGeo cg = col_geo<T>(s.g, 1);
  cg.sk = s.skip;
  cg.ncols = unsigned((x1 - x0) * s.g.Zp);
  cg.bytes = unsigned((long)(x1 - x0) * s.g.Ny * s.g.Zp * (long)sizeof(cplx<T>));
  cplx<T>* base = data + (long)x0 * s.g.Ny * s.g.Zp;
  const bool ok = with_reg_len(s.g.Ny, [&](auto lc) {
    constexpr int Lc = decltype(lc)::value;
    with_nb<T, Lc>(s.nb_half_cols, [&](auto nbc) {
      constexpr int NB = decltype(nbc)::value;
      const dim3 grid(div_up(cg.ncols, NB), narr, members), block(NB * F2<Lc>::P);
      if (fwd)
        hipLaunchKernelGGL((k_cols2<T, Lc, NB, true>), grid, block, 0, s.stream, base, stride, cg, t, mstride);
      else
        hipLaunchKernelGGL((k_cols2<T, Lc, NB, false>), grid, block, 0, s.stream, base, stride, cg, t, mstride);
      HIPCHK(hipGetLastError());
    });
  });
  REQUIRE(ok, "x-slab passes need register-FFT lengths");
}

// x-planes per slab of the fused y / z / y section (0: whole arrays, three separate passes): the planes of
// nconv arrays of all members that fit slab_bytes, when the register kernels handle y and z.
template <typename T>
inline int slab_planes(const cdft_system_s& s, int nconv) {
  if (s.slab_bytes <= 0 || !rows_reg_ok<T>(s) || !reg_fft_ok<T>(s, s.g.Ny)) return 0;
  const long per_plane = (long)nconv * s.nb * s.g.Ny * s.g.Zp * (long)sizeof(cplx<T>);
  const long p = s.slab_bytes / per_plane;
  if (p < 1 || p >= s.g.Nx) return 0;
  return (int)p;
}

// x or y pass of narr arrays (stride apart) of each of `members` batch members
// (mstride apart).  Batches need the register kernels.
template <typename T>
inline void launch_cols(const cdft_system_s& s, cplx<T>* data, long stride, int narr, int axis, bool fwd,
                        int members = 1, long mstride = 0) {
  if (launch_cols_reg<T>(s, data, stride, narr, axis, fwd, members, mstride)) return;
  REQUIRE(members == 1, "batched systems need transform lengths the register FFT supports");
  const Tables<T>& tw = tables<T>(s);
  const Radices& rad = axis == 0 ? s.radX : s.radY;
  const cplx<T>* t = (axis == 0 ? tw.x : tw.y).template as<cplx<T>>();
  const int TW = s.TWcol;
  dim3 grid(div_up(s.g.Zh, TW), axis == 0 ? s.g.Ny : s.g.Nx, narr);
  const size_t smem = sizeof(cplx<T>) * rad.L * TW;
  if (fwd)
    hipLaunchKernelGGL((k_cols<T, 16, true>), grid, dim3(kThreads), smem, s.stream, data, stride, s.g, axis, rad, t, TW);
  else
    hipLaunchKernelGGL((k_cols<T, 16, false>), grid, dim3(kThreads), smem, s.stream, data, stride, s.g, axis, rad, t,
                       TW);
  HIPCHK(hipGetLastError());
}

}  // namespace cdft_impl

namespace cdft_impl {
// Defined in cdft_api.hip.
void hess_device(cdft_system_s& s, const double* rho, const double* v, double* out, bool with_dft, bool ideal,
                 bool accumulate, cplx<double>* A2 = nullptr, cplx<double>* B2 = nullptr);
void enqueue_eval(cdft_system s, const double* rho_d, double* dF_d, bool with_dft, bool acc);
// Species::beginForceCalculation on device data: rho rescaled in place to the fixed mass (no-op without one).
void constraints_pre(cdft_system_s& s, double* rho);
// nb > 1: a batch of nb same-shape members with their own mus[nb] and V_ext
// (nb x N or none); desc->mu / desc->vext are then ignored.
// allow_trace = false keeps all 11 tensor-model arrays (the members of a mixture, which sum them per measure).
void create_system(const cdft_system_desc* desc, cdft_system* out, int nb = 1, const double* mus = nullptr,
                   const double* vexts = nullptr, bool allow_trace = true);
// Hard-sphere mixtures (cdft_api.hip): the evaluation on device data (dF written, or added to with acc; parts =
// {F, F_id, F_hs, 0, F_ext}; with_dft adds the ideal gas unless no_ideal, V_ext and -mu N per species), the number
// of species and member system i (whose stream the whole mixture uses).
int mixture_eval(cdft_mixture_s& x, const double* const* rho, double* const* dF, bool acc, double* parts,
                 bool with_dft = false, bool no_ideal = false);
int mixture_species(const cdft_mixture_s& x);
cdft_system mixture_member(const cdft_mixture_s& x, int i);
int fetch_parts(cdft_system s, double* parts);
double csum(std::initializer_list<double> v);
template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const CdftError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return CDFT_HIP_ERROR;
  }
}
}  // namespace cdft_impl
