// C ABI of the MI355X hot path (declared in include/cdft.h).
//
// A cdft_system owns: the resident weight spectra (one real number per k for
// each of the m FMT weights and the mean-field kernel), the spectral work
// arrays, the twiddle tables and a HIP stream.  cdft_system_eval runs the
// fused pipeline documented in kernels.hpp; no C++ exception crosses the ABI.
#include <cstring>
#include <mutex>
#include <random>

#include "system.hpp"
#include "pw3.hpp"
#include "mixture.hpp"

namespace cdft {
template <typename T>
long launch_pw3(const PwArgs& a, int M);  // pointwise.hip
}

namespace cdft_impl {

thread_local std::string g_err;

// Work memory of the allocation-free entry point (cdft_system_hessian_dot_noalloc): each GPU the library runs on
// holds one copy of this array until the process ends.  Host code never uses its name as an address: the device
// address of each GPU's copy comes from hipGetSymbolAddress, taken once per device at system set-up.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Records the address of the current device's scratch (set-up; the device must be current).
void scratch_setup(int dev) {
  REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return;
  void* p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  REQUIRE(p != nullptr, "no device address for the scratch array");
  g_scratch_ptr[dev] = p;
}
void* scratch_of(int dev) {
  REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  REQUIRE(g_scratch_ptr[dev], "scratch of this device not set up");
  return g_scratch_ptr[dev];
}
__global__ void k_zero_int(int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// ------------------------------- set-up ------------------------------------
void validate(const cdft_system_desc& d) {
  REQUIRE(d.Nx > 0 && d.Ny > 0 && d.Nz > 0, "grid dimensions must be positive");
  REQUIRE(d.Nz % 2 == 0, "Nz must be even");
  REQUIRE(smooth23(d.Nx) && smooth23(d.Ny) && smooth23(d.Nz / 2), "Nx, Ny and Nz/2 must be products of 2 and 3");
  REQUIRE(d.Nx <= 512 && d.Ny <= 512 && d.Nz <= 512, "grid dimensions are limited to 512");
  REQUIRE(d.dx > 0 && d.dy > 0 && d.dz > 0, "lattice spacings must be positive");
  REQUIRE(d.precision == CDFT_FP64 || d.precision == CDFT_MIXED, "unknown precision");
  REQUIRE(d.model >= CDFT_MODEL_NONE && d.model <= CDFT_WHITEBEAR_II, "unknown FMT model");
  if (d.model != CDFT_MODEL_NONE) {
    REQUIRE(d.hsd > 0, "hsd must be positive");
    // src/FMT_Species.cpp:321-322
    REQUIRE(d.fmt_weights || d.hsd >= std::min(std::min(d.dx, d.dy), d.dz), "hsd is less than the lattice spacing");
  }
  if (d.mean_field) REQUIRE(d.mf_kernel || (d.kT > 0 && d.lj_sigma > 0 && d.lj_rcut > 0), "bad mean-field parameters");
  if (d.mean_field && !d.mf_kernel) {
    REQUIRE(d.mf_potential >= CDFT_POT_LJ && d.mf_potential <= CDFT_POT_WHDF, "unknown mean-field potential");
    REQUIRE(d.mf_scheme >= CDFT_MF_QF && d.mf_scheme <= CDFT_MF_GAUSS_F, "unknown mean-field kernel rule");
    REQUIRE((d.mf_scheme != CDFT_MF_GAUSS_E && d.mf_scheme != CDFT_MF_GAUSS_F) ||
                (d.mf_ngauss >= 1 && d.mf_ngauss <= kMaxRule),
            "Gauss kernel rules need 1..32 points per axis");
    REQUIRE(d.mf_potential != CDFT_POT_WHDF || d.lj_rcut > d.lj_sigma, "WHDF needs rcut > sigma");
  }
  REQUIRE(d.weights_mode == CDFT_WEIGHTS_ACCURATE || d.weights_mode == CDFT_WEIGHTS_REFERENCE, "unknown weights mode");
  REQUIRE(d.ao_force == CDFT_AO_FORCE_EXACT || d.ao_force == CDFT_AO_FORCE_REFERENCE, "unknown AO force mode");
  REQUIRE(d.fixed_mass >= 0, "the fixed mass must be non-negative");
  REQUIRE(d.boundary >= 0 && d.boundary <= (CDFT_FIXED_BACKGROUND | CDFT_HOMOGENEOUS_BOUNDARY), "unknown boundary flags");
  // src/Species.cpp:75-76
  REQUIRE(!(d.fixed_mass > 0 && (d.boundary & CDFT_FIXED_BACKGROUND)), "Cannot have both fixed background and fixed mass");
}

void validate_ao(const cdft_system_desc& d) {
  if (!(d.ao_Rp > 0)) return;
  REQUIRE(d.model != CDFT_MODEL_NONE, "the AO polymer needs an FMT functional");
  REQUIRE(d.precision == CDFT_FP64, "the AO polymer is implemented in FP64");
  REQUIRE(d.ao_reservoir >= 0, "the AO reservoir density must be non-negative");
  REQUIRE(d.ao_weights || 2 * d.ao_Rp >= std::min(std::min(d.dx, d.dy), d.dz),
          "polymer diameter is less than the lattice spacing");
}

Pot pot_from(const cdft_potential_desc& p) {
  REQUIRE(p.kind >= CDFT_POT_LJ && p.kind <= CDFT_POT_WHDF, "unknown potential");
  return make_pot(p.kind, p.sigma, p.eps, p.rcut, p.alpha, p.bh, p.r_att_min);
}

// The system's mean-field kernel w_att (N reals, / kT) into out on the system's stream.
void launch_mf_kernel(cdft_system_s& s, double* out) {
  const cdft_system_desc& d = s.d;
  const Pot v = make_pot(d.mf_potential, d.lj_sigma, d.lj_eps, d.lj_rcut, d.mf_alpha, d.mf_bh, d.mf_r_att_min);
  const MFRule rule = make_rule(d.mf_scheme, d.mf_ngauss);
  // offsets |i| <= 1 + int(Rc / dx) + 1 per axis (src/Interaction.cpp:194-196, 272-274)
  const int lx = 2 + int(v.rcut / d.dx), ly = 2 + int(v.rcut / d.dy), lz = 2 + int(v.rcut / d.dz);
  hipLaunchKernelGGL(k_mf_kernel, dim3(div_up(s.N, kThreads)), dim3(kThreads), 0, s.stream, s.g.Nx, s.g.Ny, s.g.Nz,
                     d.dx, d.dy, d.dz, lx, ly, lz, v, rule, d.kT, out);
  HIPCHK(hipGetLastError());
}

// The 11 real-space cell weights of spheres of diameter hsd (11 * N doubles, device) on the system's stream:
// the double-double device generator, or with CDFT_WEIGHTS_REFERENCE the reference's expressions on the host.
void cell_weights(cdft_system_s& s, double hsd, double* wreal) {
  const long N = s.N;
  if (s.d.weights_mode == CDFT_WEIGHTS_REFERENCE) {
    std::vector<double> h(11 * N);
    reference_weights(s.g.Nx, s.g.Ny, s.g.Nz, s.d.dx, s.d.dy, s.d.dz, hsd, h.data());
    HIPCHK(hipMemcpyAsync(wreal, h.data(), sizeof(double) * 11 * N, hipMemcpyHostToDevice, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    return;
  }
  const double R = hsd / 2;
  const int Sxm = 2 + int(R / s.d.dx), Sym = 2 + int(R / s.d.dy), Szm = 2 + int(R / s.d.dz);
  REQUIRE(Sxm < 4 * s.g.Nx && Sym < 4 * s.g.Ny && Szm < 4 * s.g.Nz, "sphere is too large for the periodic box");
  DevBuf cells;
  const long ncell = (long)(Sxm + 1) * (Sym + 1) * (Szm + 1);
  cells.alloc(sizeof(double) * 11 * ncell);
  hipLaunchKernelGGL(k_fmt_cells, dim3(div_up(11 * ncell, 64)), dim3(64), 0, s.stream, Sxm, Sym, Szm, s.d.dx, s.d.dy,
                     s.d.dz, R, cells.as<double>());
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_fmt_gather, dim3(div_up(N, kThreads)), dim3(kThreads), 0, s.stream, s.g.Nx, s.g.Ny, s.g.Nz,
                     Sxm, Sym, Szm, cells.as<double>(), wreal);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s.stream));
}

// Per block: max |w_s - w_Txx - w_Tyy - w_Tzz| and max |w_s| over the real-space cell weights (11 arrays of n).
__global__ __launch_bounds__(256) void k_trace_dev(const double* __restrict__ w, long n, double* __restrict__ out) {
  double dev = 0.0, mx = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double ws = w[n + i];
    dev = fmax(dev, fabs(ws - w[5 * n + i] - w[8 * n + i] - w[10 * n + i]));
    mx = fmax(mx, fabs(ws));
  }
  __shared__ double sd[256], sm[256];
  sd[threadIdx.x] = dev;
  sm[threadIdx.x] = mx;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + off]);
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = sd[0];
    out[2 * blockIdx.x + 1] = sm[0];
  }
}

// out = the k = 0 value of a weight spectrum times `scale` (zero for an odd weight), in FP64.
__global__ void k_spec0(const cplx<double>* __restrict__ S, double* __restrict__ out, int odd, double scale) {
  if (threadIdx.x == 0) *out = odd ? 0.0 : S[0].x * scale;
}

// Mean split of the mixed path, one thread per member: from ms[0] = sum rho, the means of the weighted densities
// nbar_a = sum rho w_a(0) (w0 carries the spectra's 1/N; the mean-field convolution last), dPhi_a(nbar), and the
// FMT force of the means dV N sum_a dPhi_a(nbar) w_a(0).  A uniform density is exactly its mean, so the FP32 arrays
// carry only the deviations from it.
// trace: the weight set without T_zz (n_Tzz = n_s - n_Txx - n_Tyy, its dPhi folded into s, Txx and Tyy as in the
// pointwise kernels; sum_a dPhi_a w_a(0) is unchanged because w_Tzz(0) = w_s(0) - w_Txx(0) - w_Tyy(0)).
template <int MODEL>
__global__ void k_msplit(double* __restrict__ ms, const double* __restrict__ w0, int m, int nconv, PhiParams pp,
                         double dVN, int nb, int trace) {
  const int mb = blockIdx.x * blockDim.x + threadIdx.x;
  if (mb >= nb) return;
  double* q = ms + (long)mb * kMS;
  const double S = q[0];
  double n[11], out[11];
  for (int a = 0; a < 11; ++a) n[a] = a < m ? S * w0[a] : 0.0;
  if (trace) n[10] = n[1] - n[5] - n[8];
  for (int a = 0; a < nconv; ++a) q[kMSn + a] = S * w0[a];
  phi_site<MODEL>(n, pp, out);
  if (trace) {
    out[1] += out[10];
    out[5] -= out[10];
    out[8] -= out[10];
  }
  CSum K{0.0, 0.0};
  for (int a = 0; a < m; ++a) {
    q[kMSc + a] = out[a];
    K.add(out[a] * w0[a]);
  }
  q[kMSk] = (K.s + K.c) * dVN;
}

// Whether the 11 real-space weights satisfy tr w_T = w_s to rounding (1e-13 of max |w_s|): true for the lattice
// weights of FMT_Species::generateWeights (the tensor integrand's trace is the surface integrand), false for
// weights carrying the reference generator's rounding noise.
bool trace_identity_holds(cdft_system_s& s, const double* wreal) {
  const int nblk = 64;
  DevBuf out;
  out.alloc(sizeof(double) * 2 * nblk);
  hipLaunchKernelGGL(k_trace_dev, dim3(nblk), dim3(256), 0, s.stream, wreal, s.N, out.as<double>());
  HIPCHK(hipGetLastError());
  double h[2 * nblk];
  HIPCHK(hipMemcpyAsync(h, out.p, sizeof(h), hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  double dev = 0.0, mx = 0.0;
  for (int b = 0; b < nblk; ++b) {
    dev = std::max(dev, h[2 * b]);
    mx = std::max(mx, h[2 * b + 1]);
  }
  return mx > 0.0 && dev <= 1e-13 * mx;
}

// The m weight spectra (x 1/N) of spheres of diameter hsd into W: from host cell weights or generated on the device.
// With allow_trace (the colloid weights of a tensor model run by the register kernels), a weight set that satisfies
// the trace identity is stored without T_zz and s.m, s.nconv, s.trace are updated.
template <typename T>
void fmt_weight_spectra(cdft_system_s& s, double hsd, const double* host_weights, T* W, bool allow_trace = false,
                        double* w0 = nullptr) {
  const long N = s.N, Nkp = s.Nkp;
  DevBuf S;  // one double-precision spectrum
  S.alloc(sizeof(cplx<double>) * Nkp);
  const int nb = div_up(Nkp, kThreads);
  {
    DevBuf wreal;
    wreal.alloc(sizeof(double) * 11 * N);
    if (host_weights) {
      HIPCHK(hipMemcpyAsync(wreal.p, host_weights, sizeof(double) * 11 * N, hipMemcpyHostToDevice, s.stream));
    } else {
      cell_weights(s, hsd, wreal.as<double>());
    }
    HIPCHK(hipStreamSynchronize(s.stream));
    if (allow_trace && s.m == 11 && trace_identity_holds(s, wreal.as<double>())) {
      s.m = 10;
      s.nconv = s.m + (s.d.mean_field ? 1 : 0);
      s.trace = 1;
    }
    for (int a = 0; a < s.m; ++a) {
      launch_rows_r2c<double, double>(s, wreal.as<double>() + a * N, 0, S.as<cplx<double>>(), 0, 1);
      launch_cols<double>(s, S.as<cplx<double>>(), 0, 1, 1, true);
      launch_cols<double>(s, S.as<cplx<double>>(), 0, 1, 0, true);
      hipLaunchKernelGGL((k_extract_weight<T>), dim3(nb), dim3(kThreads), 0, s.stream, S.as<cplx<double>>(),
                         W + a * Nkp, Nkp, (int)((s.odd_mask >> a) & 1u), 1.0 / (double)N);
      if (w0)
        hipLaunchKernelGGL(k_spec0, dim3(1), dim3(64), 0, s.stream, S.as<cplx<double>>(), w0 + a,
                           (int)((s.odd_mask >> a) & 1u), 1.0 / (double)N);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s.stream));
  }
}

// Folds W (ColGeo::wq) when every array has the parities of the lattice weights (checked in x and in y), to
// rounding: half the weight footprint of the two x-pass convolutions.  Only with the register x pass (the LDS kernels read the full
// layout) and without the AO polymer.
template <typename T>
void fold_weight_spectra(cdft_system_s& s) {
  if (s.nconv == 0 || s.has_ao || !reg_fft_ok<T>(s, s.g.Nx)) return;
  if (const char* e = std::getenv("CDFT_NO_WFOLD"); e && e[0] == '1') return;
  const Geo& g = s.g;
  const int nblk = 64;
  DevBuf out;
  out.alloc(sizeof(double) * 2 * nblk * s.nconv);
  hipLaunchKernelGGL((k_wsym<T>), dim3(nblk, s.nconv), dim3(256), 0, s.stream, s.W.as<T>(), s.Nkp, g.Nx, g.Ny, g.Zh,
                     g.Zp, kOddX, kOddY, out.as<double>());
  HIPCHK(hipGetLastError());
  std::vector<double> h(2 * nblk * s.nconv);
  HIPCHK(hipMemcpyAsync(h.data(), out.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  const double tol = std::is_same<T, float>::value ? 1e-5 : 1e-11;
  for (int a = 0; a < s.nconv; ++a) {
    double dev = 0.0, mx = 0.0;
    for (int b = 0; b < nblk; ++b) {
      dev = std::max(dev, h[2 * (a * nblk + b)]);
      mx = std::max(mx, h[2 * (a * nblk + b) + 1]);
    }
    if (dev > tol * mx) return;  // e.g. caller weights without the lattice symmetries: keep the full layout
  }
  const long nq = (long)(g.Nx / 2 + 1) * g.Ny * g.Zp;
  DevBuf Wq;
  Wq.alloc(sizeof(T) * nq * s.nconv);
  hipLaunchKernelGGL((k_wfold<T>), dim3((unsigned)std::min<long>(div_up(nq * s.nconv, 256), 65536)), dim3(256), 0,
                     s.stream, s.W.as<T>(), Wq.as<T>(), s.Nkp, nq, s.nconv);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s.stream));
  std::swap(s.W.p, Wq.p);  // the full layout is released with Wq
  std::swap(s.W.bytes, Wq.bytes);
  s.wq = true;
  s.wq_stride = unsigned(nq);
}

// col_geo of the x pass of a convolution with the weight spectra W (folded when W is the system's folded W).
template <typename T>
ColGeo conv_geo(const cdft_system_s& s, const void* W) {
  ColGeo cg = col_geo<T>(s.g, 0);
  if (s.wq && W == s.W.p) {
    cg.wq = 1;
    cg.wstride = s.wq_stride;
    cg.pxm = kOddX;
  }
  return cg;
}

double mf_scale(const cdft_system_desc& d) { return d.mf_scale != 0 ? d.mf_scale : 1.0; }

template <typename T>
void build_weight_spectra(cdft_system_s& s) {
  const long N = s.N, Nkp = s.Nkp;
  DevBuf S;  // one double-precision spectrum
  S.alloc(sizeof(cplx<double>) * Nkp);
  const int nb = div_up(Nkp, kThreads);
  const char* nt = std::getenv("CDFT_NO_TRACE");
  const bool allow_trace = s.trace_ok && !s.has_ao && !(nt && nt[0] == '1') && rows_reg_ok<T>(s) && reg_fft_ok<T>(s, s.g.Nx) &&
                           reg_fft_ok<T>(s, s.g.Ny) && rows_reg_ok<double>(s) && reg_fft_ok<double>(s, s.g.Nx) &&
                           reg_fft_ok<double>(s, s.g.Ny);
  if (s.m > 0)
    fmt_weight_spectra<T>(s, s.d.hsd, s.d.fmt_weights, s.W.as<T>(), allow_trace,
                          s.mean_split ? s.w0.as<double>() : nullptr);
  if (s.has_ao) fmt_weight_spectra<T>(s, 2 * s.d.ao_Rp, s.d.ao_weights, s.Wao.as<T>());
  if (s.d.mean_field) {
    DevBuf w;
    w.alloc(sizeof(double) * N);
    if (s.d.mf_kernel) {
      HIPCHK(hipMemcpyAsync(w.p, s.d.mf_kernel, sizeof(double) * N, hipMemcpyHostToDevice, s.stream));
    } else {
      launch_mf_kernel(s, w.as<double>());
      HIPCHK(hipGetLastError());
    }
    // a_vdw = sum_R w_att(R) (src/Interaction.cpp:290-296, 42-46), times scale_interaction's factor
    const int nblk = 256;
    hipLaunchKernelGGL(k_sum_real, dim3(nblk), dim3(kThreads), 0, s.stream, w.as<double>(), N,
                       s.partials.as<double2>());
    hipLaunchKernelGGL(k_reduce_partials_wide, dim3(1), dim3(1024), 0, s.stream, s.partials.as<double2>(), (long)nblk, 1,
                       s.results.as<double>(), 0L, 0L);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&s.a_vdw, s.results.p, sizeof(double), hipMemcpyDeviceToHost, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    s.a_vdw *= mf_scale(s.d);
    launch_rows_r2c<double, double>(s, w.as<double>(), 0, S.as<cplx<double>>(), 0, 1);
    launch_cols<double>(s, S.as<cplx<double>>(), 0, 1, 1, true);
    launch_cols<double>(s, S.as<cplx<double>>(), 0, 1, 0, true);
    // v = (dV/N) c2r(rho_k w_k): fold dV/N into the stored spectrum
    const double dV = s.d.dx * s.d.dy * s.d.dz;
    hipLaunchKernelGGL((k_extract_weight<T>), dim3(nb), dim3(kThreads), 0, s.stream, S.as<cplx<double>>(),
                       s.W.as<T>() + s.m * Nkp, Nkp, 0, mf_scale(s.d) * dV / (double)N);
    if (s.mean_split)
      hipLaunchKernelGGL(k_spec0, dim3(1), dim3(64), 0, s.stream, S.as<cplx<double>>(), s.w0.as<double>() + s.m, 0,
                         mf_scale(s.d) * dV / (double)N);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s.stream));
  }
  fold_weight_spectra<T>(s);
}

void create_system(const cdft_system_desc* desc, cdft_system* out, int nb, const double* mus, const double* vexts,
                   bool allow_trace) {
  REQUIRE(desc && out, "null argument");
  REQUIRE(nb >= 1, "member count must be positive");
  validate(*desc);
  auto s = std::make_unique<cdft_system_s>();
  s->d = *desc;
  s->nb = nb;
  s->trace_ok = allow_trace;
  HIPCHK(hipGetDevice(&s->device));
  scratch_setup(s->device);
  HIPCHK(hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking));
  s->stream = s->own_stream;
  Geo& g = s->g;
  g.Nx = desc->Nx;
  g.Ny = desc->Ny;
  g.Nz = desc->Nz;
  g.Zh = desc->Nz / 2 + 1;
  g.Zp = (g.Zh + 7) / 8 * 8;
  s->N = (long)g.Nx * g.Ny * g.Nz;
  s->Nkp = (long)g.Nx * g.Ny * g.Zp;
  s->nrows = (long)g.Nx * g.Ny;
  s->mixed = desc->precision == CDFT_MIXED;
  s->m = desc->model == CDFT_MODEL_NONE ? 0 : (desc->model >= CDFT_ESFMT ? 11 : 5);
  s->nconv = s->m + (desc->mean_field ? 1 : 0);
  s->radX = make_radices(g.Nx);
  s->radY = make_radices(g.Ny);
  s->radM = make_radices(std::max(1, g.Nz / 2));
  const int M = g.Nz / 2;
  s->RBrow = std::max(1, 1024 / std::max(1, M));
  s->RBpw = std::max(1, 256 / std::max(1, M));
  const int Lmax = std::max(g.Nx, g.Ny);
  s->TWcol = std::max(1, std::min(16, 4096 / Lmax));
  const int esz = s->mixed ? 8 : 16;
  s->TWconv = std::max(1, std::min(16, (65536 / (2 * esz)) / Lmax));
  if (const char* e = std::getenv("CDFT_LDS_FFT")) s->reg_fft = !(e[0] == '1');
  // measured on MI355X at 256^3: FP64 x/y passes and backward convolution run
  // faster with 8 columns per block than with 16 (occupancy vs coalescing)
  s->nb_half_cols = s->nb_half_bwd = !s->mixed;
  if (const char* e = std::getenv("CDFT_NB_HALF")) {  // three digits: x/y passes, forward, backward
    const size_t n = std::strlen(e);
    if (n > 0) s->nb_half_cols = e[0] == '1';
    if (n > 1) s->nb_half_fwd = e[1] == '1';
    if (n > 2) s->nb_half_bwd = e[2] == '1';
  }
  if (const char* e = std::getenv("CDFT_PW_ROWS")) s->pw_rows = std::atoi(e);
  if (const char* e = std::getenv("CDFT_PW3")) s->pw3 = e[0] != '0';
  // y / z / y section in x-slabs of ~192 MB (all arrays of all members): each slab stays in the 256 MB
  // last-level cache between its three kernels (measured optimum at 256^3, FP64 and mixed)
  s->slab_bytes = 192L << 20;
  if (const char* e = std::getenv("CDFT_SLAB_MB")) s->slab_bytes = std::atol(e) << 20;
  s->twd.build(g);
  if (s->mixed) s->twf.build(g);
  const size_t cs = s->mixed ? sizeof(cplx<float>) : sizeof(cplx<double>);
  const size_t rs = s->mixed ? sizeof(float) : sizeof(double);
  REQUIRE(nb == 1 || (rows_reg_ok<double>(*s) && reg_fft_ok<double>(*s, g.Nx) && reg_fft_ok<double>(*s, g.Ny) &&
                      s->Nkp * (long)cs * s->nconv * nb < (1L << 40)),
          "batched systems need Nx, Ny and Nz/2 in {4,6,8,12,16,24,32,48,64,96,128,256}");
  s->A.alloc(cs * s->Nkp * nb);
  if (s->nconv) {
    s->B.alloc(cs * s->Nkp * s->nconv * nb);
    s->W.alloc(rs * s->Nkp * s->nconv);
  }
  if (desc->mean_field) s->vmf.alloc(sizeof(double) * s->N * nb);
  validate_ao(*desc);
  if (desc->ao_Rp > 0) {
    s->has_ao = true;
    s->Wao.alloc(rs * s->Nkp * s->m);
    s->psi.alloc(sizeof(double) * s->N * nb);
    s->A2.alloc(cs * s->Nkp * nb);
    s->B2.alloc(cs * s->Nkp * s->nconv * nb);
  }
  s->rho.alloc(sizeof(double) * s->N * nb);
  s->dF.alloc(sizeof(double) * s->N * nb);
  s->pslots = std::max<long>(s->nrows, 256);  // at least one row per block in every z kernel
  s->partials.alloc(sizeof(double2) * s->pslots * nb);
  s->partials2.alloc(sizeof(double2) * s->pslots * 4 * nb);
  s->flag.alloc(sizeof(int) * nb);
  HIPCHK(hipMemset(s->flag.p, 0, sizeof(int) * nb));
  s->results.alloc(sizeof(double) * 8 * nb);
  s->lagrange.assign(nb, mus ? 0.0 : desc->mu);
  if (mus) for (int b = 0; b < nb; ++b) s->lagrange[b] = mus[b];
  if (desc->fixed_mass > 0) {
    s->masses.assign(nb, desc->fixed_mass);
    s->masses_d.alloc(sizeof(double) * nb);
    HIPCHK(hipMemcpy(s->masses_d.p, s->masses.data(), sizeof(double) * nb, hipMemcpyHostToDevice));
  }
  const double* vx = vexts ? vexts : desc->vext;
  if (vx) {
    s->vext.alloc(sizeof(double) * s->N * nb);
    HIPCHK(hipMemcpy(s->vext.p, vx, sizeof(double) * s->N * nb, hipMemcpyHostToDevice));
    s->has_vext = true;
  }
  REQUIRE(nb == 1 || mus, "batched systems need one mu per member");
  if (mus) {
    s->mus.alloc(sizeof(double) * nb);
    HIPCHK(hipMemcpy(s->mus.p, mus, sizeof(double) * nb, hipMemcpyHostToDevice));
    if (nb == 1) s->d.mu = mus[0];
  }
  // mean split in the mixed path: needs the register kernels (x convolution, pointwise and final z passes)
  const char* nms = std::getenv("CDFT_NO_MEAN_SPLIT");
  s->mean_split = s->mixed && s->m > 0 && !(nms && nms[0] == '1') && rows_reg_ok<float>(*s) &&
                  reg_fft_ok<float>(*s, g.Nx);
  if (s->mean_split) {
    s->msplit.alloc(sizeof(double) * kMS * nb);
    s->w0.alloc(sizeof(double) * 12);
  }
  if (s->mixed) build_weight_spectra<float>(*s);
  else build_weight_spectra<double>(*s);
  // host pointers of the descriptor are not retained
  s->d.fmt_weights = nullptr;
  s->d.vext = nullptr;
  s->d.mf_kernel = nullptr;
  s->d.ao_weights = nullptr;
  s->ev.resize(64);  // per-slab stage marks; cdft_system_get_timing sums them per stage name
  s->ev_names.resize(64);
  for (auto& e : s->ev) HIPCHK(hipEventCreate(&e));
  *out = s.release();
}

// Pointwise z pass over the rows of x-planes [x0, x1) (all planes by default); its partial sums go to block
// slots blk_off, blk_off + 1, ...  Returns the number of blocks.
template <typename T, int MODEL>
long launch_pointwise(cdft_system_s& s, int nconv, int x0 = 0, int x1 = -1, long blk_off = 0) {
  const Tables<T>& tw = tables<T>(s);
  const double* ms = std::is_same<T, float>::value && s.mean_split ? s.msplit.as<double>() : nullptr;
  const int M = s.g.Nz / 2;
  if (x1 < 0) x1 = s.g.Nx;
  const long nrows = (long)(x1 - x0) * s.g.Ny;
  const PhiParams pp = make_phi_params(s.d.hsd, s.d.A, s.d.B, MODEL == M_WBI || MODEL == M_WBII);
  if (rows_reg_ok<T>(s) && s.pw3 && s.pw_rows <= 0) {
    const PwArgs pa{s.B.as<cplx<T>>() + (long)x0 * s.g.Ny * s.g.Zp, s.Nkp, s.m, nconv,
                    s.vmf.p ? s.vmf.as<double>() + (long)x0 * s.g.Ny * s.g.Nz : nullptr, nrows, s.g.Zp,
                    tw.mp.p, tw.n.p, pp, s.partials.as<double2>() + blk_off, s.flag.as<int>(),
                    Batch{s.Nkp, s.nconv * s.Nkp, s.N, s.pslots, nullptr, s.skip}, s.trace, MODEL, s.nb, s.stream,
                    ms};
    const long nblk = launch_pw3<T>(pa, M);
    REQUIRE(nblk != -2, "launch of the pointwise kernel failed");
    if (nblk > 0) return nblk;
  }
  if (rows_reg_ok<T>(s)) {
    long nblk = -1;
    with_reg_len(M, [&](auto mc) {
      constexpr int Mc = decltype(mc)::value;
      using RF = RowF<T, Mc>;
      const int per = nconv * RF::SLOT * int(sizeof(cplx<T>));
      int RB = std::min(kThreads / (nconv * RF::P), (65536 - 4096) / per);
      if (s.pw_rows > 0) RB = std::min(RB, s.pw_rows);
      if (RB < 1) return;
      nblk = div_up(nrows, RB);
      const Batch bt{s.Nkp, s.nconv * s.Nkp, s.N, s.pslots, nullptr, s.skip};
      double* vmf = s.vmf.p ? s.vmf.as<double>() + (long)x0 * s.g.Ny * s.g.Nz : nullptr;
      hipLaunchKernelGGL((k_z_pointwise2<T, MODEL, Mc>), dim3(nblk, s.nb), dim3(nconv * RB * RF::P),
                         size_t(RB) * per, s.stream, s.B.as<cplx<T>>() + (long)x0 * s.g.Ny * s.g.Zp, s.Nkp, s.m, nconv,
                         vmf, nrows, s.g.Zp, tw.mp.template as<cplx<T>>(), tw.n.template as<cplx<T>>(), RB, pp,
                         s.partials.as<double2>() + blk_off, s.flag.as<int>(), bt, s.trace, ms);
      HIPCHK(hipGetLastError());
    });
    if (nblk > 0) return nblk;
  }
  REQUIRE(x0 == 0 && x1 == s.g.Nx, "x-slab passes need the register z kernels");
  REQUIRE(s.nb == 1, "batched systems need transform lengths the register FFT supports");
  REQUIRE(!s.trace, "the trace-reduced weight set needs the register z kernels");
  const int RB = s.RBpw;
  const size_t smem = sizeof(cplx<T>) * nconv * RB * M;
  hipLaunchKernelGGL((k_z_pointwise<T, MODEL, 16>), dim3(div_up(s.nrows, RB)), dim3(kThreads), smem, s.stream,
                     s.B.as<cplx<T>>(), s.Nkp, s.m, nconv, s.vmf.as<double>(), s.g, s.radM, tw.m.template as<cplx<T>>(),
                     tw.n.template as<cplx<T>>(), RB, pp, s.partials.as<double2>(), s.flag.as<int>());
  HIPCHK(hipGetLastError());
  return div_up(s.nrows, RB);
}

// The full evaluation on device pointers.  with_dft = false: FMT only.
// mu of the full evaluation: zero with a fixed mass (Species::beginForceCalculation, src/Species.cpp:64-71)
double eval_mu(const cdft_system_s& s) { return s.d.fixed_mass > 0 ? 0.0 : s.d.mu; }

// ---- species constraints (constraints.hpp) ----
dim3 cons_grid(const cdft_system_s& s, long n) {
  return dim3((unsigned)std::max<long>(1, std::min<long>(div_up(n, kThreads), 4096)), s.nb);
}
// one compensated sum per member into results[8 b + slot]
void cons_sum(cdft_system_s& s, const double* a, const double* b, long n, int mode, int slot) {
  const long nblk = std::min<long>(s.pslots, std::max<long>(1, std::min<long>(div_up(n, kThreads), 1024)));
  hipLaunchKernelGGL(k_cons_sum, dim3((unsigned)nblk, s.nb), dim3(kThreads), 0, s.stream, a, b, n, s.N, mode, s.g.Nx,
                     s.g.Ny, s.g.Nz, s.partials.as<double2>(), s.pslots);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_partials_wide, dim3(s.nb), dim3(1024), 0, s.stream, s.partials.as<double2>(), nblk, 1,
                     s.results.as<double>() + slot, s.pslots, 8L);
  HIPCHK(hipGetLastError());
}
// Species::beginForceCalculation: rho rescaled in place to the fixed mass
void constraints_pre(cdft_system_s& s, double* rho) {
  if (!(s.d.fixed_mass > 0)) return;
  const double dV = s.d.dx * s.d.dy * s.d.dz;
  cons_sum(s, rho, nullptr, s.N, CONS_SUM, 6);
  hipLaunchKernelGGL(k_cons_scale, cons_grid(s, s.N), dim3(kThreads), 0, s.stream, rho, s.N, s.N,
                     s.results.as<double>() + 6, 8L, s.masses_d.as<double>(), dV);
  HIPCHK(hipGetLastError());
}
// Species::endForceCalculation: fixed background, homogeneous boundary, fixed-mass projection (in that order)
void constraints_post(cdft_system_s& s, const double* rho, double* dF) {
  const long nbnd = boundary_count(s.g.Nx, s.g.Ny, s.g.Nz);
  if (s.d.boundary & CDFT_FIXED_BACKGROUND) {
    hipLaunchKernelGGL(k_cons_bset, cons_grid(s, nbnd), dim3(kThreads), 0, s.stream, dF, nbnd, s.N, s.g.Nx, s.g.Ny,
                       s.g.Nz, nullptr, 8L, 0.0);
    HIPCHK(hipGetLastError());
  }
  if (s.d.boundary & CDFT_HOMOGENEOUS_BOUNDARY) {
    cons_sum(s, dF, nullptr, nbnd, CONS_BSUM, 6);
    hipLaunchKernelGGL(k_cons_bset, cons_grid(s, nbnd), dim3(kThreads), 0, s.stream, dF, nbnd, s.N, s.g.Nx, s.g.Ny,
                       s.g.Nz, s.results.as<double>() + 6, 8L, 1.0 / (double)nbnd);
    HIPCHK(hipGetLastError());
  }
  if (s.d.fixed_mass > 0) {
    cons_sum(s, dF, rho, s.N, CONS_DOT, 7);
    hipLaunchKernelGGL(k_cons_shift, cons_grid(s, s.N), dim3(kThreads), 0, s.stream, dF, s.N, s.N,
                       s.results.as<double>() + 7, 8L, s.masses_d.as<double>(), s.d.dx * s.d.dy * s.d.dz);
    HIPCHK(hipGetLastError());
  }
}

template <typename T>
void eval_device(cdft_system_s& s, const double* rho, double* dF, bool with_dft, bool accumulate) {
  s.n_marks = 0;
  s.mark("start");
  const long Nkp = s.Nkp;
  const int nconv = with_dft ? s.nconv : s.m;
  const bool mf = with_dft && s.d.mean_field;
  cplx<T>* A = s.A.as<cplx<T>>();
  cplx<T>* B = s.B.as<cplx<T>>();
  const Tables<T>& tw = tables<T>(s);
  const int nb = s.nb;
  const long sA = Nkp, sB = s.nconv * Nkp;  // member strides of A and of the weight-array block B
  const bool msplit = std::is_same<T, float>::value && s.mean_split;
  long zblk = 0;
  if (nconv > 0) {
    // the z pass of rho; in the mixed path it also sums rho per member in FP64 (the mean split)
    launch_rows_r2c<T, double>(s, rho, s.N, A, sA, nb, msplit ? s.partials.as<double2>() : nullptr, s.pslots,
                               &zblk);
    s.mark("z_r2c");
  }
  if (nconv > 0 && msplit) {
    // the means of the weighted densities and dPhi at them (k_msplit)
    hipLaunchKernelGGL(k_reduce_partials_wide, dim3(nb), dim3(1024), 0, s.stream, s.partials.as<double2>(), zblk, 1,
                       s.msplit.as<double>(), s.pslots, (long)kMS);
    const PhiParams pp = make_phi_params(s.d.hsd, s.d.A, s.d.B,
                                         s.d.model == CDFT_WHITEBEAR_I || s.d.model == CDFT_WHITEBEAR_II);
    const double dVN = s.d.dx * s.d.dy * s.d.dz * (double)s.N;
    const dim3 g1(div_up(nb, 64)), b1(64);
    double* msp = s.msplit.as<double>();
    const double* w0 = s.w0.as<double>();
    switch (s.d.model) {
      case CDFT_ROSENFELD: hipLaunchKernelGGL(k_msplit<M_ROSENFELD>, g1, b1, 0, s.stream, msp, w0, s.m, nconv, pp, dVN, nb, s.trace); break;
      case CDFT_RSLT: hipLaunchKernelGGL(k_msplit<M_RSLT>, g1, b1, 0, s.stream, msp, w0, s.m, nconv, pp, dVN, nb, s.trace); break;
      case CDFT_ESFMT: hipLaunchKernelGGL(k_msplit<M_ESFMT>, g1, b1, 0, s.stream, msp, w0, s.m, nconv, pp, dVN, nb, s.trace); break;
      case CDFT_WHITEBEAR_I: hipLaunchKernelGGL(k_msplit<M_WBI>, g1, b1, 0, s.stream, msp, w0, s.m, nconv, pp, dVN, nb, s.trace); break;
      default: hipLaunchKernelGGL(k_msplit<M_WBII>, g1, b1, 0, s.stream, msp, w0, s.m, nconv, pp, dVN, nb, s.trace); break;
    }
    HIPCHK(hipGetLastError());
    s.mark("mean_split");
  }
  if (nconv > 0) {
    launch_cols<T>(s, A, 0, 1, 1, true, nb, sA);
    s.mark("y_fwd");
    if (reg_fft_ok<T>(s, s.g.Nx)) {
      const ColGeo cg = conv_geo<T>(s, s.W.p);
      with_reg_len(s.g.Nx, [&](auto lc) {
        constexpr int Lc = decltype(lc)::value;
        with_nb<T, Lc>(s.nb_half_fwd, [&](auto nbc) {
          constexpr int NB = decltype(nbc)::value;
          const Batch bt{sA, sB, s.N, 0, nullptr, s.skip};
          hipLaunchKernelGGL((k_conv_fwd2<T, Lc, NB>), dim3(div_up(cg.ncols, NB), nb), dim3(NB * F2<Lc>::P), 0,
                             s.stream, A, B, Nkp, s.W.as<T>(), nconv, s.odd_mask, cg, tw.x.template as<cplx<T>>(), bt,
                             (int)msplit);
          HIPCHK(hipGetLastError());
        });
      });
    } else {
      REQUIRE(nb == 1, "batched systems need transform lengths the register FFT supports");
      REQUIRE(!s.wq, "folded weight spectra are read by the register x pass only");
      const int TW = s.TWconv;
      dim3 grid(div_up(s.g.Zh, TW), s.g.Ny);
      hipLaunchKernelGGL((k_conv_fwd<T, 16>), grid, dim3(kThreads), 2 * sizeof(cplx<T>) * s.g.Nx * TW, s.stream, A, B,
                         Nkp, s.W.as<T>(), nconv, s.odd_mask, s.g, s.radX, tw.x.template as<cplx<T>>(), TW);
      HIPCHK(hipGetLastError());
    }
    s.mark("x_conv_fwd");
    const int planes = s.m > 0 ? slab_planes<T>(s, nconv) : 0;
    auto pointwise = [&](int x0, int x1, long off) -> long {
      switch (s.d.model) {
        case CDFT_ROSENFELD: return launch_pointwise<T, M_ROSENFELD>(s, nconv, x0, x1, off);
        case CDFT_RSLT: return launch_pointwise<T, M_RSLT>(s, nconv, x0, x1, off);
        case CDFT_ESFMT: return launch_pointwise<T, M_ESFMT>(s, nconv, x0, x1, off);
        case CDFT_WHITEBEAR_I: return launch_pointwise<T, M_WBI>(s, nconv, x0, x1, off);
        default: return launch_pointwise<T, M_WBII>(s, nconv, x0, x1, off);
      }
    };
    if (planes > 0) {
      // y inverse, pointwise z and y forward x-slab by x-slab: a slab's arrays stay in the last-level cache
      // between the three kernels instead of making three round trips through HBM
      long nblk = 0;
      for (int x0 = 0; x0 < s.g.Nx; x0 += planes) {
        const int x1 = std::min(s.g.Nx, x0 + planes);
        launch_cols_slab<T>(s, B, Nkp, nconv, false, nb, sB, x0, x1);
        s.mark("y_inv");
        nblk += pointwise(x0, x1, nblk);
        s.mark("z_pointwise");
        launch_cols_slab<T>(s, B, Nkp, s.m, true, nb, sB, x0, x1);
        s.mark("y_fwd2");
      }
      hipLaunchKernelGGL(k_reduce_partials_wide, dim3(nb), dim3(1024), 0, s.stream, s.partials.as<double2>(), nblk, 1,
                         s.results.as<double>(), s.pslots, 8L);
      s.mark("z_pointwise");
    } else {
      launch_cols<T>(s, B, Nkp, nconv, 1, false, nb, sB);
      s.mark("y_inv");
    }
    if (s.m > 0 && planes == 0) {
      const long nblk = pointwise(0, s.g.Nx, 0);
      hipLaunchKernelGGL(k_reduce_partials_wide, dim3(nb), dim3(1024), 0, s.stream, s.partials.as<double2>(), nblk, 1,
                         s.results.as<double>(), s.pslots, 8L);
      s.mark("z_pointwise");
      launch_cols<T>(s, B, Nkp, s.m, 1, true, nb, sB);
      s.mark("y_fwd2");
    }
    if (s.m > 0) {
      if (reg_fft_ok<T>(s, s.g.Nx)) {
        const ColGeo cg = conv_geo<T>(s, s.W.p);
        with_reg_len(s.g.Nx, [&](auto lc) {
          constexpr int Lc = decltype(lc)::value;
          with_nb<T, Lc>(s.nb_half_bwd, [&](auto nbc) {
            constexpr int NB = decltype(nbc)::value;
            const Batch bt{sA, sB, s.N, 0, nullptr, s.skip};
            hipLaunchKernelGGL((k_conv_bwd2<T, Lc, NB>), dim3(div_up(cg.ncols, NB), nb), dim3(NB * F2<Lc>::P), 0,
                               s.stream, B, A, Nkp, s.W.as<T>(), s.m, s.odd_mask, cg, tw.x.template as<cplx<T>>(),
                               bt);
            HIPCHK(hipGetLastError());
          });
        });
      } else {
        REQUIRE(!s.wq, "folded weight spectra are read by the register x pass only");
        const int TW = s.TWconv;
        dim3 grid(div_up(s.g.Zh, TW), s.g.Ny);
        hipLaunchKernelGGL((k_conv_bwd<T, 16>), grid, dim3(kThreads), 2 * sizeof(cplx<T>) * s.g.Nx * TW, s.stream, B,
                           A, Nkp, s.W.as<T>(), s.m, s.odd_mask, s.g, s.radX, tw.x.template as<cplx<T>>(), TW);
        HIPCHK(hipGetLastError());
      }
      s.mark("x_conv_bwd");
      launch_cols<T>(s, A, 0, 1, 1, false, nb, sA);
      s.mark("y_inv2");
    } else {
      // mean field only: inverse z of the single convolved array -> vmf
      launch_rows_c2r<T, double>(s, B, sB, s.vmf.as<double>(), s.N, nb, 1.0);
      s.mark("z_c2r_mf");
    }
  }
  FinalParams fp{};
  fp.dV = s.d.dx * s.d.dy * s.d.dz;
  fp.mu = with_dft ? eval_mu(s) : 0.0;
  fp.ideal = with_dft ? (s.d.ideal && !s.no_ideal) : 0;
  fp.fmt = s.m > 0;
  fp.mf = mf;
  fp.ext = with_dft && s.has_vext;
  fp.accumulate = accumulate;
  fp.ms = msplit && s.m > 0 ? s.msplit.as<double>() : nullptr;
  {
    const int M = s.g.Nz / 2;
    long nblk = 0;
    if (rows_reg_ok<T>(s)) {
      with_reg_len(M, [&](auto mc) {
        constexpr int Mc = decltype(mc)::value;
        constexpr int RB = row_rb<T, Mc>(32768), P = RowF<T, Mc>::P, SLOT = RowF<T, Mc>::SLOT;
        nblk = div_up(s.nrows, RB);
        const Batch bt{sA, sB, s.N, 4 * s.pslots,
                       with_dft && s.nb > 1 && !(s.d.fixed_mass > 0) ? s.mus.as<double>() : nullptr, s.skip};
        hipLaunchKernelGGL((k_z_final2<T, Mc>), dim3(nblk, nb), dim3(RB * P), RB * SLOT * sizeof(cplx<T>), s.stream, A,
                           rho, s.vext.as<double>(), s.vmf.as<double>(), dF, unsigned(s.N * 8), s.nrows, s.g.Zp,
                           tw.m.template as<cplx<T>>(), tw.n.template as<cplx<T>>(), RB, fp,
                           s.partials2.as<double2>(), bt);
      });
    } else {
      REQUIRE(nb == 1, "batched systems need transform lengths the register FFT supports");
      const int RB = s.RBrow;
      nblk = div_up(s.nrows, RB);
      hipLaunchKernelGGL((k_z_final<T, 4>), dim3(nblk), dim3(kThreads), sizeof(cplx<T>) * RB * M, s.stream, A, rho,
                         s.vext.as<double>(), s.vmf.as<double>(), dF, s.g, s.radM, tw.m.template as<cplx<T>>(),
                         tw.n.template as<cplx<T>>(), RB, fp, s.partials2.as<double2>());
    }
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_partials_wide, dim3(nb), dim3(1024), 0, s.stream, s.partials2.as<double2>(), nblk, 4,
                       s.results.as<double>() + 1, 4 * s.pslots, 8L);
    HIPCHK(hipGetLastError());
  }
  s.mark("z_final");
}

// Two-sum based accurate addition of a few terms.
double csum(std::initializer_list<double> v) {
  double s = 0.0, c = 0.0;
  for (double x : v) {
    const double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  return s + c;
}

// Enqueue one evaluation on device pointers (no synchronisation).
// Forward half of the convolution pipeline on FP64 device data: in (nb members of N reals) -> A (spectra) ->
// B (nconv weighted-density arrays, x and y inverse-transformed, z still spectral).  Register-FFT lengths only.
void forward_weighted(cdft_system_s& s, const double* in, cplx<double>* A, cplx<double>* B, int nconv,
                      const double* W = nullptr) {
  if (!W) W = s.W.as<double>();
  const Tables<double>& tw = tables<double>(s);
  const long Nkp = s.Nkp, sA = Nkp, sB = s.nconv * Nkp;
  launch_rows_r2c<double, double>(s, in, s.N, A, sA, s.nb);
  launch_cols<double>(s, A, 0, 1, 1, true, s.nb, sA);
  const ColGeo cg = conv_geo<double>(s, W);
  with_reg_len(s.g.Nx, [&](auto lc) {
    constexpr int Lc = decltype(lc)::value;
    with_nb<double, Lc>(s.nb_half_fwd, [&](auto nbc) {
      constexpr int NB = decltype(nbc)::value;
      const Batch bt{sA, sB, s.N, 0, nullptr};
      hipLaunchKernelGGL((k_conv_fwd2<double, Lc, NB>), dim3(div_up(cg.ncols, NB), s.nb), dim3(NB * F2<Lc>::P), 0,
                         s.stream, A, B, Nkp, W, nconv, s.odd_mask, cg, tw.x.template as<cplx<double>>(), bt, 0);
      HIPCHK(hipGetLastError());
    });
  });
  launch_cols<double>(s, B, Nkp, nconv, 1, false, s.nb, sB);
}

template <int MODEL>
void launch_hess(cdft_system_s& s, bool mf, double inv_dp = 0.0, double scale = 1.0,
                 const cplx<double>* B2 = nullptr, int single_offdiag = 0) {
  if (!B2) B2 = s.B2.as<cplx<double>>();
  const Tables<double>& tw = tables<double>(s);
  const int M = s.g.Nz / 2;
  const PhiParams pp = make_phi_params(s.d.hsd, s.d.A, s.d.B, MODEL == M_WBI || MODEL == M_WBII);
  bool done = false;
  with_reg_len(M, [&](auto mc) {
    constexpr int Mc = decltype(mc)::value;
    using RF = RowF<double, Mc>;
    const int nseq = 2 * s.m + (mf ? 1 : 0);
    const int per = nseq * RF::SLOT * int(sizeof(cplx<double>));
    const int RB = std::min(512 / (nseq * RF::P), (65536 - 4096) / per);
    if (RB < 1) return;
    const Batch bt{s.Nkp, s.nconv * s.Nkp, s.N, 0, nullptr};
    hipLaunchKernelGGL((k_z_hess2<MODEL, Mc>), dim3(div_up(s.nrows, RB), s.nb), dim3(nseq * RB * RF::P),
                       size_t(RB) * per, s.stream, s.B.as<cplx<double>>(), B2, s.Nkp, s.m,
                       mf ? 1 : 0, s.vmf.as<double>(), s.nrows, s.g.Zp, tw.mp.template as<cplx<double>>(),
                       tw.n.template as<cplx<double>>(), RB, pp, inv_dp > 0 ? inv_dp : pp.inv_d, scale,
                       s.flag.as<int>(), bt, s.trace, single_offdiag);
    HIPCHK(hipGetLastError());
    done = true;
  });
  REQUIRE(done, "grid too long in z for the Hessian-vector product (needs 2m+1 rows of Nz in 64 KB of LDS)");
}

// Backward convolution (x pass, sum_a w_a D_a, inverse x) of m FP64 arrays D into Acc with weight spectra W.
void conv_bwd_f64(cdft_system_s& s, const cplx<double>* D, cplx<double>* Acc, const double* W, int m) {
  const Tables<double>& tw = tables<double>(s);
  const long Nkp = s.Nkp, sA = Nkp, sB = s.nconv * Nkp;
  const ColGeo cg = conv_geo<double>(s, W);
  with_reg_len(s.g.Nx, [&](auto lc) {
    constexpr int Lc = decltype(lc)::value;
    with_nb<double, Lc>(s.nb_half_bwd, [&](auto nbc) {
      constexpr int NB = decltype(nbc)::value;
      const Batch bt{sA, sB, s.N, 0, nullptr};
      hipLaunchKernelGGL((k_conv_bwd2<double, Lc, NB>), dim3(div_up(cg.ncols, NB), s.nb), dim3(NB * F2<Lc>::P), 0,
                         s.stream, D, Acc, Nkp, W, m, s.odd_mask, cg, tw.x.template as<cplx<double>>(), bt);
      HIPCHK(hipGetLastError());
    });
  });
}

template <int MODEL>
long launch_pw_ao(cdft_system_s& s, int nconv) {
  const Tables<double>& tw = tables<double>(s);
  const int M = s.g.Nz / 2;
  const PhiParams pp = make_phi_params(s.d.hsd, s.d.A, s.d.B, MODEL == M_WBI || MODEL == M_WBII);
  long nblk = -1;
  with_reg_len(M, [&](auto mc) {
    constexpr int Mc = decltype(mc)::value;
    using RF = RowF<double, Mc>;
    const int nseq = nconv + s.m;
    const int per = nseq * RF::SLOT * int(sizeof(cplx<double>));
    const int RB = std::min(512 / (nseq * RF::P), (65536 - 4096) / per);
    if (RB < 1) return;
    nblk = div_up(s.nrows, RB);
    const Batch bt{s.Nkp, s.nconv * s.Nkp, s.N, s.pslots, nullptr};
    hipLaunchKernelGGL((k_z_pw_ao<MODEL, Mc>), dim3(nblk, s.nb), dim3(nseq * RB * RF::P), size_t(RB) * per, s.stream,
                       s.B.as<cplx<double>>(), s.B2.as<cplx<double>>(), s.Nkp, s.m, nconv, s.vmf.as<double>(), s.nrows,
                       s.g.Zp, tw.mp.template as<cplx<double>>(), tw.n.template as<cplx<double>>(), RB, pp,
                       1.0 / (2 * s.d.ao_Rp), s.partials.as<double2>(), s.flag.as<int>(), bt);
    HIPCHK(hipGetLastError());
  });
  REQUIRE(nblk > 0, "grid too long in z for the AO functional (needs 2m+1 rows of Nz in 64 KB of LDS)");
  return nblk;
}

// FMT::calculateFreeEnergyAndDerivatives with an FMT_AO_Species (src/FMT.cpp:363-427, src/FMT_Species.cpp:647-738)
// plus the rest of the DFT, FP64: (1) n of rho; (2) Phi, dPhi collapsed for the colloid and the polymer; (3) the
// colloid back-convolution; (4) psi = sum_a w^p_a * dPhi^p_a; (5) exp(-psi) and F_AO = rho_p sum (1 - exp(-psi));
// (6) Upsilon = w^p * exp(-psi); (7) n again; (8) rho_p d2Phi(n) . Upsilon (polymer measures) collapsed for the
// colloid; (9) its back-convolution; (10) the final pass adds the colloid and AO forces.  The tensor off-diagonals
// of (8) are counted twice, as the derivative of F_AO requires.
void eval_device_ao(cdft_system_s& s, const double* rho, double* dF, bool with_dft, bool accumulate) {
  REQUIRE(rows_reg_ok<double>(s) && reg_fft_ok<double>(s, s.g.Nx) && reg_fft_ok<double>(s, s.g.Ny),
          "the AO polymer needs register-FFT transform lengths");
  s.n_marks = 0;
  s.mark("start");
  const int nconv = with_dft ? s.nconv : s.m;
  const bool mf = with_dft && s.d.mean_field;
  const long Nkp = s.Nkp, sA = Nkp, sB = s.nconv * Nkp;
  const int nb = s.nb;
  cplx<double>* A = s.A.as<cplx<double>>();
  cplx<double>* B = s.B.as<cplx<double>>();
  cplx<double>* A2 = s.A2.as<cplx<double>>();
  cplx<double>* B2 = s.B2.as<cplx<double>>();
  const Tables<double>& tw = tables<double>(s);
  forward_weighted(s, rho, A, B, nconv);
  s.mark("forward");
  long nblk = 0;
  switch (s.d.model) {
    case CDFT_ROSENFELD: nblk = launch_pw_ao<M_ROSENFELD>(s, nconv); break;
    case CDFT_RSLT: nblk = launch_pw_ao<M_RSLT>(s, nconv); break;
    case CDFT_ESFMT: nblk = launch_pw_ao<M_ESFMT>(s, nconv); break;
    case CDFT_WHITEBEAR_I: nblk = launch_pw_ao<M_WBI>(s, nconv); break;
    default: nblk = launch_pw_ao<M_WBII>(s, nconv); break;
  }
  hipLaunchKernelGGL(k_reduce_partials_wide, dim3(nb), dim3(1024), 0, s.stream, s.partials.as<double2>(), nblk, 1,
                     s.results.as<double>(), s.pslots, 8L);
  s.mark("z_pointwise_ao");
  launch_cols<double>(s, B, Nkp, s.m, 1, true, nb, sB);
  conv_bwd_f64(s, B, A, s.W.as<double>(), s.m);
  launch_cols<double>(s, A, 0, 1, 1, false, nb, sA);
  s.mark("colloid_back");
  launch_cols<double>(s, B2, Nkp, s.m, 1, true, nb, sB);
  conv_bwd_f64(s, B2, A2, s.Wao.as<double>(), s.m);
  launch_cols<double>(s, A2, 0, 1, 1, false, nb, sA);
  launch_rows_c2r<double, double>(s, A2, sA, s.psi.as<double>(), s.N, nb, 1.0);
  {
    const int eb = 256;
    hipLaunchKernelGGL(k_ao_exp, dim3(eb, nb), dim3(kThreads), 0, s.stream, s.psi.as<double>(), s.N, s.N,
                       s.partials.as<double2>(), s.pslots);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_partials_wide, dim3(nb), dim3(1024), 0, s.stream, s.partials.as<double2>(), (long)eb, 1,
                       s.results.as<double>() + 5, s.pslots, 8L);
  }
  s.mark("psi");
  forward_weighted(s, s.psi.as<double>(), A2, B2, s.m, s.Wao.as<double>());
  forward_weighted(s, rho, A2, B, s.m);
  s.mark("upsilon");
  const int so = s.d.ao_force == CDFT_AO_FORCE_REFERENCE;
  switch (s.d.model) {
    case CDFT_ROSENFELD: launch_hess<M_ROSENFELD>(s, false, 1.0 / (2 * s.d.ao_Rp), s.d.ao_reservoir, nullptr, so); break;
    case CDFT_RSLT: launch_hess<M_RSLT>(s, false, 1.0 / (2 * s.d.ao_Rp), s.d.ao_reservoir, nullptr, so); break;
    case CDFT_ESFMT: launch_hess<M_ESFMT>(s, false, 1.0 / (2 * s.d.ao_Rp), s.d.ao_reservoir, nullptr, so); break;
    case CDFT_WHITEBEAR_I: launch_hess<M_WBI>(s, false, 1.0 / (2 * s.d.ao_Rp), s.d.ao_reservoir, nullptr, so); break;
    default: launch_hess<M_WBII>(s, false, 1.0 / (2 * s.d.ao_Rp), s.d.ao_reservoir, nullptr, so); break;
  }
  launch_cols<double>(s, B, Nkp, s.m, 1, true, nb, sB);
  conv_bwd_f64(s, B, A2, s.W.as<double>(), s.m);
  launch_cols<double>(s, A2, 0, 1, 1, false, nb, sA);
  s.mark("ao_force");
  const int M = s.g.Nz / 2;
  for (int pass = 0; pass < 2; ++pass) {
    FinalParams fp{};
    fp.dV = s.d.dx * s.d.dy * s.d.dz;
    fp.fmt = 1;
    if (pass == 0) {
      fp.mu = with_dft ? eval_mu(s) : 0.0;
      fp.ideal = with_dft ? (s.d.ideal && !s.no_ideal) : 0;
      fp.mf = mf;
      fp.ext = with_dft && s.has_vext;
      fp.accumulate = accumulate;
    } else {
      fp.accumulate = 1;  // add dV c2r(AO force spectrum)
    }
    long fblk = 0;
    with_reg_len(M, [&](auto mc) {
      constexpr int Mc = decltype(mc)::value;
      constexpr int RB = row_rb<double, Mc>(32768), P = RowF<double, Mc>::P, SLOT = RowF<double, Mc>::SLOT;
      fblk = div_up(s.nrows, RB);
      const Batch bt{sA, sB, s.N, 4 * s.pslots,
                     pass == 0 && with_dft && s.nb > 1 && !(s.d.fixed_mass > 0) ? s.mus.as<double>() : nullptr};
      hipLaunchKernelGGL((k_z_final2<double, Mc>), dim3(fblk, nb), dim3(RB * P), RB * SLOT * sizeof(cplx<double>),
                         s.stream, pass == 0 ? A : A2, rho, s.vext.as<double>(), s.vmf.as<double>(), dF,
                         unsigned(s.N * 8), s.nrows, s.g.Zp, tw.m.template as<cplx<double>>(),
                         tw.n.template as<cplx<double>>(), RB, fp, s.partials2.as<double2>(), bt);
    });
    HIPCHK(hipGetLastError());
    if (pass == 0)
      hipLaunchKernelGGL(k_reduce_partials_wide, dim3(nb), dim3(1024), 0, s.stream, s.partials2.as<double2>(), fblk, 4,
                         s.results.as<double>() + 1, 4 * s.pslots, 8L);
  }
  HIPCHK(hipGetLastError());
  s.mark("z_final");
}

// out = [out] + (d2 beta Omega / drho drho) v on FP64 device data (with_dft = false: FMT part only).  The spectra
// of the direction v go to A2 (nb spectra) and B2 (nconv nb spectra): the system's buffers (allocated on first use)
// unless the caller passes its own.
void hess_device(cdft_system_s& s, const double* rho, const double* v, double* out, bool with_dft, bool ideal,
                 bool accumulate, cplx<double>* A2, cplx<double>* B2) {
  REQUIRE(!s.mixed, "the Hessian-vector product needs an FP64 system");
  REQUIRE(!s.has_ao, "the Hessian-vector product is not implemented with an AO polymer");
  REQUIRE(rows_reg_ok<double>(s) && reg_fft_ok<double>(s, s.g.Nx) && reg_fft_ok<double>(s, s.g.Ny),
          "the Hessian-vector product needs register-FFT transform lengths");
  const bool mf = with_dft && s.d.mean_field;
  const long Nkp = s.Nkp, sA = Nkp, sB = s.nconv * Nkp;
  const size_t cs = sizeof(cplx<double>);
  if (!A2 || !B2) {
    if (s.A2.bytes < cs * Nkp * s.nb) s.A2.alloc(cs * Nkp * s.nb);
    if (s.B2.bytes < cs * Nkp * s.nconv * s.nb) s.B2.alloc(cs * Nkp * s.nconv * s.nb);
    A2 = s.A2.as<cplx<double>>();
    B2 = s.B2.as<cplx<double>>();
  }
  cplx<double>* A = s.A.as<cplx<double>>();
  cplx<double>* B = s.B.as<cplx<double>>();
  const Tables<double>& tw = tables<double>(s);
  if (s.m > 0) forward_weighted(s, rho, A, B, s.m);
  if (s.m > 0 || mf) forward_weighted(s, v, A2, B2, s.m + (mf ? 1 : 0));
  if (s.m > 0) {
    switch (s.d.model) {
      case CDFT_ROSENFELD: launch_hess<M_ROSENFELD>(s, mf, 0.0, 1.0, B2); break;
      case CDFT_RSLT: launch_hess<M_RSLT>(s, mf, 0.0, 1.0, B2); break;
      case CDFT_ESFMT: launch_hess<M_ESFMT>(s, mf, 0.0, 1.0, B2); break;
      case CDFT_WHITEBEAR_I: launch_hess<M_WBI>(s, mf, 0.0, 1.0, B2); break;
      default: launch_hess<M_WBII>(s, mf, 0.0, 1.0, B2); break;
    }
    launch_cols<double>(s, B, Nkp, s.m, 1, true, s.nb, sB);
    const ColGeo cg = conv_geo<double>(s, s.W.p);
    with_reg_len(s.g.Nx, [&](auto lc) {
      constexpr int Lc = decltype(lc)::value;
      with_nb<double, Lc>(s.nb_half_bwd, [&](auto nbc) {
        constexpr int NB = decltype(nbc)::value;
        const Batch bt{sA, sB, s.N, 0, nullptr};
        hipLaunchKernelGGL((k_conv_bwd2<double, Lc, NB>), dim3(div_up(cg.ncols, NB), s.nb), dim3(NB * F2<Lc>::P), 0,
                           s.stream, B, A, Nkp, s.W.as<double>(), s.m, s.odd_mask, cg,
                           tw.x.template as<cplx<double>>(), bt);
        HIPCHK(hipGetLastError());
      });
    });
    launch_cols<double>(s, A, 0, 1, 1, false, s.nb, sA);
  } else if (mf) {
    launch_rows_c2r<double, double>(s, B2, sB, s.vmf.as<double>(), s.N, s.nb, 1.0);
  }
  FinalParams fp{};
  fp.dV = s.d.dx * s.d.dy * s.d.dz;
  fp.ideal = with_dft && ideal && s.d.ideal;
  fp.fmt = s.m > 0;
  fp.mf = mf;
  fp.accumulate = accumulate;
  fp.hess = 1;
  const int M = s.g.Nz / 2;
  with_reg_len(M, [&](auto mc) {
    constexpr int Mc = decltype(mc)::value;
    constexpr int RB = row_rb<double, Mc>(32768), P = RowF<double, Mc>::P, SLOT = RowF<double, Mc>::SLOT;
    const long nblk = div_up(s.nrows, RB);
    const Batch bt{sA, sB, s.N, 4 * s.pslots, nullptr};
    hipLaunchKernelGGL((k_z_final2<double, Mc>), dim3(nblk, s.nb), dim3(RB * P), RB * SLOT * sizeof(cplx<double>),
                       s.stream, A, rho, v, s.vmf.as<double>(), out, unsigned(s.N * 8), s.nrows, s.g.Zp,
                       tw.m.template as<cplx<double>>(), tw.n.template as<cplx<double>>(), RB, fp,
                       s.partials2.as<double2>(), bt);
  });
  HIPCHK(hipGetLastError());
}

// Bulk AO polymer terms (FMT::BulkFex / BulkMuex, src/FMT.cpp:616-716): arg = sum_a dPhi_a(n) w^p_a is the
// polymer's insertion free energy, pref = d arg / d rho (the second derivative of Phi along the colloid weights,
// from the dual part); f_AO = rho_p (1 - e^-arg), mu_AO = rho_p pref e^-arg.
template <int MODEL>
void bulk_ao(const double* om, const double* omp, double rho, const PhiParams& pp, double dp, double* arg,
             double* pref) {
  constexpr bool TENSOR = MODEL >= M_ESFMT;
  Dual n[11], out[11];
  for (int a = 0; a < 11; ++a) n[a] = Dual(rho * om[a], om[a]);
  const double id = pp.inv_d, id2 = pp.inv_d2;
  const Dual v2[3] = {n[2], n[3], n[4]}, v1[3] = {n[2] * id, n[3] * id, n[4] * id};
  PhiDeriv<Dual> dd;
  phi_meas<MODEL, Dual>(n[0], n[1] * id2, n[1] * id, n[1], v1, v2, n + 5, pp, dd);
  phi_collapse<TENSOR, Dual>(dd, 1.0 / dp, 1.0 / (dp * dp), out);
  Dual acc(0.0, 0.0);
  for (int a = 0; a < 11; ++a) acc = acc + out[a] * omp[a];
  *arg = acc.v;
  *pref = acc.d;
}

// One evaluation on device data.  With a fixed mass the full evaluation rescales rho_d in place first
// (the species density, as Species::beginForceCalculation does).
void enqueue_eval(cdft_system s, const double* rho_d, double* dF_d, bool with_dft, bool acc) {
  if (with_dft) constraints_pre(*s, const_cast<double*>(rho_d));
  if (s->has_ao) eval_device_ao(*s, rho_d, dF_d, with_dft, acc);
  else if (s->mixed) eval_device<float>(*s, rho_d, dF_d, with_dft, acc);
  else eval_device<double>(*s, rho_d, dF_d, with_dft, acc);
  if (with_dft) constraints_post(*s, rho_d, dF_d);
  s->last_with_dft = with_dft;
  s->last_ideal = with_dft && s->d.ideal && !s->no_ideal;
  s->last_fixed_mass = with_dft && s->d.fixed_mass > 0;
}

// F parts of one member from its five device results (sum Phi, sum rho ln rho - rho, sum rho v_mf, sum V rho, sum rho).
void parts_from_results(const cdft_system_s& s, const double* res, double mu, double* parts) {
  const bool with_dft = s.last_with_dft;
  const double dV = s.d.dx * s.d.dy * s.d.dz;
  // F_AO = rho_p sum (1 - exp(-psi)) (src/FMT_Species.cpp:712-724) joins the FMT sum before the factor dV
  const double F_hs = s.m > 0 ? (s.has_ao ? csum({res[0], s.d.ao_reservoir * res[5]}) : res[0]) * dV : 0.0;
  const double F_id = s.last_ideal ? res[1] : 0.0;
  const double F_mf = (with_dft && s.d.mean_field) ? 0.5 * res[2] : 0.0;
  const double mu_ = with_dft ? mu : 0.0;
  const double F_ext = (with_dft && s.has_vext ? res[3] * dV : 0.0) - res[4] * dV * mu_;
  parts[0] = csum({F_id, F_hs, F_mf, F_ext});
  parts[1] = F_id;
  parts[2] = F_hs;
  parts[3] = F_mf;
  parts[4] = F_ext;
}

// Wait for the stream and turn the device results into the F parts (5 per
// member).  CDFT_ETA_TOO_LARGE if any member's flag is set (flags cleared).
int fetch_parts(cdft_system s, double* parts) {
  const int nb = s->nb;
  std::vector<double> res(8 * (size_t)nb), mus(nb, eval_mu(*s));
  std::vector<int> fl(nb, 0);
  HIPCHK(hipMemcpyAsync(res.data(), s->results.p, sizeof(double) * 8 * nb, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(fl.data(), s->flag.p, sizeof(int) * nb, hipMemcpyDeviceToHost, s->stream));
  if (nb > 1 && !(s->d.fixed_mass > 0))
    HIPCHK(hipMemcpyAsync(mus.data(), s->mus.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  bool bad = false;
  for (int b = 0; b < nb; ++b) bad = bad || fl[b];
  if (bad) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(kThreads), 0, s->stream, s->flag.as<int>(), nb);
    HIPCHK(hipStreamSynchronize(s->stream));
    g_err = "Eta too large in FMT::dPHI";
    return CDFT_ETA_TOO_LARGE;
  }
  if (s->last_fixed_mass)  // mu_ = sum dF rho / fixed mass (src/Species.cpp:99-110)
    for (int b = 0; b < nb; ++b) s->lagrange[b] = res[8 * b + 7] / s->masses[b];
  if (parts)
    for (int b = 0; b < nb; ++b) parts_from_results(*s, res.data() + 8 * b, mus[b], parts + 5 * b);
  return CDFT_OK;
}

int run_eval(cdft_system s, const double* rho, double* dF, double* parts, int flags, bool with_dft) {
  REQUIRE(s && rho && dF, "null argument");
  const bool dev = flags & CDFT_DEVICE_PTRS;
  const bool acc = flags & CDFT_ACCUMULATE;
  const size_t bytes = sizeof(double) * s->N * s->nb;
  const double* rho_d = rho;
  double* dF_d = dF;
  if (!dev) {
    HIPCHK(hipMemcpyAsync(s->rho.p, rho, bytes, hipMemcpyHostToDevice, s->stream));
    if (acc) HIPCHK(hipMemcpyAsync(s->dF.p, dF, bytes, hipMemcpyHostToDevice, s->stream));
    rho_d = s->rho.as<double>();
    dF_d = s->dF.as<double>();
  }
  s->no_ideal = with_dft && (flags & CDFT_NO_IDEAL);
  enqueue_eval(s, rho_d, dF_d, with_dft, acc);
  s->no_ideal = false;
  if (!dev) HIPCHK(hipMemcpyAsync(dF, s->dF.p, bytes, hipMemcpyDeviceToHost, s->stream));
  return fetch_parts(s, parts);
}

// Standalone transform helper (a system with no functional terms).
int standalone_fft(int Nx, int Ny, int Nz, int precision, const double* in, double* out, bool r2c) {
  cdft_system_desc d{};
  d.Nx = Nx;
  d.Ny = Ny;
  d.Nz = Nz;
  d.dx = d.dy = d.dz = 1.0;
  d.model = CDFT_MODEL_NONE;
  d.precision = precision;
  cdft_system s = nullptr;
  create_system(&d, &s);
  std::unique_ptr<cdft_system_s> guard(s);
  const long N = s->N;
  const int Zh = s->g.Zh;
  const long Nk = (long)Nx * Ny * Zh;
  if (r2c) {
    HIPCHK(hipMemcpy(s->rho.p, in, sizeof(double) * N, hipMemcpyHostToDevice));
    std::vector<double> host(2 * s->Nkp);
    if (s->mixed) {
      launch_rows_r2c<float, double>(*s, s->rho.as<double>(), 0, s->A.as<cplx<float>>(), 0, 1);
      launch_cols<float>(*s, s->A.as<cplx<float>>(), 0, 1, 1, true);
      launch_cols<float>(*s, s->A.as<cplx<float>>(), 0, 1, 0, true);
      std::vector<float> hf(2 * s->Nkp);
      HIPCHK(hipMemcpyAsync(hf.data(), s->A.p, sizeof(float) * 2 * s->Nkp, hipMemcpyDeviceToHost, s->stream));
      HIPCHK(hipStreamSynchronize(s->stream));
      for (size_t i = 0; i < hf.size(); ++i) host[i] = hf[i];
    } else {
      launch_rows_r2c<double, double>(*s, s->rho.as<double>(), 0, s->A.as<cplx<double>>(), 0, 1);
      launch_cols<double>(*s, s->A.as<cplx<double>>(), 0, 1, 1, true);
      launch_cols<double>(*s, s->A.as<cplx<double>>(), 0, 1, 0, true);
      HIPCHK(hipMemcpyAsync(host.data(), s->A.p, sizeof(double) * 2 * s->Nkp, hipMemcpyDeviceToHost, s->stream));
      HIPCHK(hipStreamSynchronize(s->stream));
    }
    for (long r = 0; r < (long)Nx * Ny; ++r)
      std::memcpy(out + 2 * r * Zh, host.data() + 2 * r * s->g.Zp, sizeof(double) * 2 * Zh);
    (void)Nk;
  } else {
    std::vector<double> host(2 * s->Nkp, 0.0);
    for (long r = 0; r < (long)Nx * Ny; ++r)
      std::memcpy(host.data() + 2 * r * s->g.Zp, in + 2 * r * Zh, sizeof(double) * 2 * Zh);
    if (s->mixed) {
      std::vector<float> hf(host.begin(), host.end());
      HIPCHK(hipMemcpy(s->A.p, hf.data(), sizeof(float) * hf.size(), hipMemcpyHostToDevice));
      launch_cols<float>(*s, s->A.as<cplx<float>>(), 0, 1, 0, false);
      launch_cols<float>(*s, s->A.as<cplx<float>>(), 0, 1, 1, false);
      launch_rows_c2r<float, double>(*s, s->A.as<cplx<float>>(), 0, s->dF.as<double>(), 0, 1, 1.0);
    } else {
      HIPCHK(hipMemcpy(s->A.p, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice));
      launch_cols<double>(*s, s->A.as<cplx<double>>(), 0, 1, 0, false);
      launch_cols<double>(*s, s->A.as<cplx<double>>(), 0, 1, 1, false);
      launch_rows_c2r<double, double>(*s, s->A.as<cplx<double>>(), 0, s->dF.as<double>(), 0, 1, 1.0);
    }
    HIPCHK(hipMemcpyAsync(out, s->dF.p, sizeof(double) * N, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  return CDFT_OK;
}

// ------------------------------------------------------------------ mixtures
// The FMT of several hard-sphere species (mixture.hpp).  Each species is an FP64 FMT-only member system holding its
// weight spectra and transform buffers; all members share the first member's stream.
template <int MODEL>
long launch_pw_mix(cdft_system_s& s, cplx<double>* Mx, int nm, double2* partials, int* flag, int x0, int x1) {
  const Tables<double>& tw = tables<double>(s);
  const PhiParams pp = make_phi_params(s.d.hsd, s.d.A, s.d.B, MODEL == M_WBI || MODEL == M_WBII);
  const long nrows = (long)(x1 - x0) * s.g.Ny;  // the rows of the x-planes [x0, x1)
  cplx<double>* base = Mx + (long)x0 * s.g.Ny * s.g.Zp;
  long nblk = -1;
  with_reg_len(s.g.Nz / 2, [&](auto mc) {
    constexpr int Mc = decltype(mc)::value;
    using RF = RowF<double, Mc>;
    const int per = nm * RF::SLOT * int(sizeof(cplx<double>));
    const int RB = std::min(kThreads / (nm * RF::P), (65536 - 4096) / per);
    if (RB < 1) return;
    nblk = div_up(nrows, RB);
    hipLaunchKernelGGL((k_z_pw_mix<MODEL, Mc>), dim3(nblk), dim3(nm * RB * RF::P), size_t(RB) * per, s.stream, base,
                       s.Nkp, nm, nrows, s.g.Zp, tw.mp.template as<cplx<double>>(),
                       tw.n.template as<cplx<double>>(), RB, pp, partials, flag);
    HIPCHK(hipGetLastError());
  });
  REQUIRE(nblk > 0, "the mixture's z rows do not fit the pointwise kernel");
  return nblk;
}

}  // namespace cdft_impl

struct cdft_mixture_s {
  std::vector<cdft_system> sp;  // one member system per species
  std::vector<double> hsd;
  int nm = 0;                   // measure arrays (16 with the tensor, 10 without)
  DevBuf Mx, partials, flag, results;
  // per species: the nm forward then the nm backward measure weight spectra (mix_fwd_coef / mix_bwd_coef times the
  // species' weight spectra, in its weight layout), wlen values each
  std::vector<std::unique_ptr<DevBuf>> wm;
  std::vector<long> wlen;
  ~cdft_mixture_s() {  // the later members run on the first member's stream, destroyed with it
    for (size_t i = 1; i < sp.size(); ++i) delete sp[i];
    if (!sp.empty()) delete sp[0];
  }
};

namespace cdft_impl {

void create_mixture(const cdft_mixture_desc* d, cdft_mixture* out) {
  REQUIRE(d && out, "null argument");
  REQUIRE(d->nspecies >= 1 && d->hsd, "a mixture needs at least one species and its diameters");
  REQUIRE(d->model >= CDFT_ROSENFELD && d->model <= CDFT_WHITEBEAR_II, "a mixture needs an FMT model");
  auto x = std::make_unique<cdft_mixture_s>();
  for (int i = 0; i < d->nspecies; ++i) {
    cdft_system_desc sd{};
    sd.Nx = d->Nx;
    sd.Ny = d->Ny;
    sd.Nz = d->Nz;
    sd.dx = d->dx;
    sd.dy = d->dy;
    sd.dz = d->dz;
    sd.model = d->model;
    sd.A = d->A;
    sd.B = d->B;
    sd.hsd = d->hsd[i];
    sd.precision = CDFT_FP64;
    sd.fmt_weights = d->fmt_weights ? d->fmt_weights[i] : nullptr;
    sd.weights_mode = d->weights_mode;
    sd.ideal = d->ideal;
    sd.mu = d->mu ? d->mu[i] : 0.0;
    sd.vext = d->vext ? d->vext[i] : nullptr;
    sd.fixed_mass = d->fixed_mass ? d->fixed_mass[i] : 0.0;
    sd.boundary = d->boundary ? d->boundary[i] : 0;
    cdft_system s = nullptr;
    create_system(&sd, &s, 1, nullptr, nullptr, true);
    x->sp.push_back(s);
    x->hsd.push_back(d->hsd[i]);
    if (i > 0) {  // one stream for the whole mixture
      HIPCHK(hipStreamSynchronize(s->own_stream));
      HIPCHK(hipStreamDestroy(s->own_stream));
      s->own_stream = nullptr;
      s->stream = x->sp[0]->stream;
    }
  }
  cdft_system_s& s0 = *x->sp[0];
  REQUIRE(rows_reg_ok<double>(s0) && reg_fft_ok<double>(s0, s0.g.Nx) && reg_fft_ok<double>(s0, s0.g.Ny),
          "mixtures need Nx, Ny and Nz/2 in {4,6,8,12,16,24,32,48,64,96,128,256}");
  // tensor models: 16 measures, or 15 when every species' weights satisfy the trace identity (stored without T_zz)
  bool all_trace = true;
  for (cdft_system m : x->sp) all_trace = all_trace && m->trace;
  x->nm = s0.m >= 10 ? (all_trace ? kMixMeas - 1 : kMixMeas) : 10;
  for (int i = 0; i < d->nspecies; ++i) {
    cdft_system_s& s = *x->sp[i];
    const long len = s.wq ? (long)s.wq_stride : s.Nkp;
    auto w = std::make_unique<DevBuf>();
    w->alloc(sizeof(double) * 2 * x->nm * len);
    const dim3 g(std::min(div_up(len, 256), 4096));
    const double* W = s.W.as<double>();
    for (int b = 0; b < x->nm; ++b) {
      double* fw = w->as<double>() + b * len;
      double* bw = w->as<double>() + (x->nm + b) * len;
      if (b == 15 && s.trace) {  // T_zz measure of a species stored without its T_zz weight
        hipLaunchKernelGGL(k_mix_tzz, g, dim3(256), 0, s0.stream, W + len, W + 5 * len, W + 8 * len, fw, len,
                           mix_fwd_coef(b, d->hsd[i]));
        hipLaunchKernelGGL(k_mix_tzz, g, dim3(256), 0, s0.stream, W + len, W + 5 * len, W + 8 * len, bw, len,
                           mix_bwd_coef(b, d->hsd[i]));
        continue;
      }
      const double* src = W + mix_src(b) * len;
      hipLaunchKernelGGL(k_mix_scale, g, dim3(256), 0, s0.stream, src, fw, len, mix_fwd_coef(b, d->hsd[i]));
      hipLaunchKernelGGL(k_mix_scale, g, dim3(256), 0, s0.stream, src, bw, len, mix_bwd_coef(b, d->hsd[i]));
    }
    HIPCHK(hipGetLastError());
    x->wm.push_back(std::move(w));
    x->wlen.push_back(len);
  }
  HIPCHK(hipStreamSynchronize(s0.stream));
  x->Mx.alloc(sizeof(cplx<double>) * s0.Nkp * x->nm);
  x->partials.alloc(sizeof(double2) * s0.pslots);
  x->flag.alloc(sizeof(int));
  HIPCHK(hipMemset(x->flag.p, 0, sizeof(int)));
  x->results.alloc(sizeof(double) * 8);
  *out = x.release();
}

// The mixture on device data: dF_s written, or added to with acc; parts (may be null) = {F, F_id, F_hs, 0, F_ext}
// summed over species.  with_dft adds each species' ideal gas (unless no_ideal), external field and -mu N
// (src/DFT.cpp:377-425) to the hard-sphere term.
// The x-pass geometry of species i's measure weight spectra (its weight layout, measure parities).
ColGeo mix_geo(const cdft_mixture_s& x, int i) {
  const cdft_system_s& s = *x.sp[i];
  ColGeo cg = col_geo<double>(s.g, 0);
  if (s.wq) {
    cg.wq = 1;
    cg.wstride = (unsigned)x.wlen[i];
    cg.pxm = kMixOddX;
  }
  return cg;
}

// Species i's share of the measures: its density spectrum (A, after the z and y passes) times the forward measure
// weight spectra, inverse x, written to Mx (first species) or added to it.
void mix_conv_fwd(cdft_mixture_s& x, int i, const cplx<double>* A, cplx<double>* Mx) {
  cdft_system_s& s = *x.sp[i];
  const Tables<double>& tw = tables<double>(s);
  const ColGeo cg = mix_geo(x, i);
  const double* W = x.wm[i]->as<double>();
  with_reg_len(s.g.Nx, [&](auto lc) {
    constexpr int Lc = decltype(lc)::value;
    with_nb<double, Lc>(s.nb_half_fwd, [&](auto nbc) {
      constexpr int NB = decltype(nbc)::value;
      const Batch bt{s.Nkp, x.nm * s.Nkp, s.N, 0, nullptr};
      if (i == 0)
        hipLaunchKernelGGL((k_conv_fwd2<double, Lc, NB, false>), dim3(div_up(cg.ncols, NB), 1), dim3(NB * F2<Lc>::P),
                           0, s.stream, A, Mx, s.Nkp, W, x.nm, kMixOdd, cg, tw.x.template as<cplx<double>>(), bt, 0);
      else
        hipLaunchKernelGGL((k_conv_fwd2<double, Lc, NB, true>), dim3(div_up(cg.ncols, NB), 1), dim3(NB * F2<Lc>::P),
                           0, s.stream, A, Mx, s.Nkp, W, x.nm, kMixOdd, cg, tw.x.template as<cplx<double>>(), bt, 0);
      HIPCHK(hipGetLastError());
    });
  });
}

// Species i's force spectrum (x real): the sum over measures of the derivative spectra dM times its backward
// measure weight spectra, into Acc.
void mix_conv_bwd(cdft_mixture_s& x, int i, const cplx<double>* dM, cplx<double>* Acc) {
  cdft_system_s& s = *x.sp[i];
  const Tables<double>& tw = tables<double>(s);
  const ColGeo cg = mix_geo(x, i);
  const double* W = x.wm[i]->as<double>() + (long)x.nm * x.wlen[i];
  with_reg_len(s.g.Nx, [&](auto lc) {
    constexpr int Lc = decltype(lc)::value;
    with_nb<double, Lc>(s.nb_half_bwd, [&](auto nbc) {
      constexpr int NB = decltype(nbc)::value;
      const Batch bt{s.Nkp, x.nm * s.Nkp, s.N, 0, nullptr};
      hipLaunchKernelGGL((k_conv_bwd2<double, Lc, NB>), dim3(div_up(cg.ncols, NB), 1), dim3(NB * F2<Lc>::P), 0,
                         s.stream, dM, Acc, s.Nkp, W, x.nm, kMixOdd, cg, tw.x.template as<cplx<double>>(), bt);
      HIPCHK(hipGetLastError());
    });
  });
}

int mixture_species(const cdft_mixture_s& x) { return (int)x.sp.size(); }
cdft_system mixture_member(const cdft_mixture_s& x, int i) { return x.sp[i]; }

int mixture_eval(cdft_mixture_s& x, const double* const* rho, double* const* dF, bool acc, double* parts,
                 bool with_dft, bool no_ideal) {
  cdft_system_s& s0 = *x.sp[0];
  const long Nkp = s0.Nkp;
  const int S = (int)x.sp.size();
  cplx<double>* Mx = x.Mx.as<cplx<double>>();
  // 1. each species' density spectrum (z and y passes) times its forward measure weights, inverse x, summed into the
  // measures, which then take one inverse y pass; with the DFT terms a fixed-mass species is first rescaled to its
  // mass (Species::beginForceCalculation)
  for (int i = 0; i < S; ++i) {
    cdft_system_s& s = *x.sp[i];
    if (with_dft) constraints_pre(s, const_cast<double*>(rho[i]));
    cplx<double>* A = s.A.as<cplx<double>>();
    launch_rows_r2c<double, double>(s, rho[i], s.N, A, Nkp, 1);
    launch_cols<double>(s, A, 0, 1, 1, true, 1, Nkp);
    mix_conv_fwd(x, i, A, Mx);  // the measures summed over species in the (x, ky, kz) domain
  }
  // 2. inverse y of the measures, Phi and its measure derivatives per site (z inverse and forward in the kernel),
  // forward y of the derivatives: x-slab by x-slab when the slab of all nm arrays fits the cache budget, so that a
  // slab stays in the last-level cache between the three kernels
  const int Nx = s0.g.Nx;
  const long per_plane = (long)x.nm * s0.g.Ny * s0.g.Zp * (long)sizeof(cplx<double>);
  const long fit = s0.slab_bytes > 0 ? s0.slab_bytes / per_plane : 0;
  const int planes = fit >= 1 && fit < Nx ? (int)fit : Nx;
  auto pw = [&](int x0, int x1, double2* part) -> long {
    switch (s0.d.model) {
      case CDFT_ROSENFELD: return launch_pw_mix<M_ROSENFELD>(s0, Mx, x.nm, part, x.flag.as<int>(), x0, x1);
      case CDFT_RSLT: return launch_pw_mix<M_RSLT>(s0, Mx, x.nm, part, x.flag.as<int>(), x0, x1);
      case CDFT_ESFMT: return launch_pw_mix<M_ESFMT>(s0, Mx, x.nm, part, x.flag.as<int>(), x0, x1);
      case CDFT_WHITEBEAR_I: return launch_pw_mix<M_WBI>(s0, Mx, x.nm, part, x.flag.as<int>(), x0, x1);
      default: return launch_pw_mix<M_WBII>(s0, Mx, x.nm, part, x.flag.as<int>(), x0, x1);
    }
  };
  long nblk = 0;
  for (int x0 = 0; x0 < Nx; x0 += planes) {
    const int x1 = std::min(Nx, x0 + planes);
    if (planes < Nx) launch_cols_slab<double>(s0, Mx, Nkp, x.nm, false, 1, 0, x0, x1);
    else launch_cols<double>(s0, Mx, Nkp, x.nm, 1, false, 1, 0);
    nblk += pw(x0, x1, x.partials.as<double2>() + nblk);
    if (planes < Nx) launch_cols_slab<double>(s0, Mx, Nkp, x.nm, true, 1, 0, x0, x1);
    else launch_cols<double>(s0, Mx, Nkp, x.nm, 1, true, 1, 0);
  }
  REQUIRE(nblk <= s0.pslots, "partial-sum slots of the mixture's pointwise pass");
  hipLaunchKernelGGL(k_reduce_partials_wide, dim3(1), dim3(1024), 0, s0.stream, x.partials.as<double2>(), nblk, 1,
                     x.results.as<double>(), 0L, 8L);
  HIPCHK(hipGetLastError());
  // 3. per species: x back-convolution of the measure derivatives with its backward measure weights (the collapse
  // onto its arrays folded in), y and z inverse
  const double dV = s0.d.dx * s0.d.dy * s0.d.dz;
  for (int i = 0; i < S; ++i) {
    cdft_system_s& s = *x.sp[i];
    FinalParams fp{};
    fp.dV = dV;
    fp.fmt = 1;
    fp.accumulate = acc;
    fp.mu = with_dft ? eval_mu(s) : 0.0;
    fp.ideal = with_dft && s.d.ideal && !no_ideal;
    fp.ext = with_dft && s.has_vext;
    cplx<double>* A = s.A.as<cplx<double>>();
    mix_conv_bwd(x, i, Mx, A);  // the collapse onto this species folded into its measure weights
    launch_cols<double>(s, A, 0, 1, 1, false, 1, Nkp);
    const Tables<double>& tw = tables<double>(s);
    with_reg_len(s.g.Nz / 2, [&](auto mc) {
      constexpr int Mc = decltype(mc)::value;
      constexpr int RB = row_rb<double, Mc>(32768), P = RowF<double, Mc>::P, SLOT = RowF<double, Mc>::SLOT;
      const Batch bt{Nkp, s.nconv * Nkp, s.N, 4 * s.pslots, nullptr};
      const long fblk = div_up(s.nrows, RB);
      hipLaunchKernelGGL((k_z_final2<double, Mc>), dim3(fblk, 1), dim3(RB * P), RB * SLOT * sizeof(cplx<double>),
                         s0.stream, A, rho[i], s.vext.as<double>(), nullptr, dF[i], unsigned(s.N * 8), s.nrows,
                         s.g.Zp, tw.m.template as<cplx<double>>(), tw.n.template as<cplx<double>>(), RB, fp,
                         s.partials2.as<double2>(), bt);
      HIPCHK(hipGetLastError());
      // this species' ideal-gas, external-field and particle sums into its results[1..4]
      hipLaunchKernelGGL(k_reduce_partials_wide, dim3(1), dim3(1024), 0, s0.stream, s.partials2.as<double2>(), fblk,
                         4, s.results.as<double>() + 1, 4 * s.pslots, 8L);
    });
    HIPCHK(hipGetLastError());
    // Species::endForceCalculation on this species' whole force (boundaries, fixed-mass projection)
    if (with_dft) constraints_post(s, rho[i], dF[i]);
    s.last_fixed_mass = with_dft && s.d.fixed_mass > 0;
  }
  double res = 0.0;
  int bad = 0;
  std::vector<double> sres(8 * (size_t)S);
  HIPCHK(hipMemcpyAsync(&res, x.results.p, sizeof(double), hipMemcpyDeviceToHost, s0.stream));
  HIPCHK(hipMemcpyAsync(&bad, x.flag.p, sizeof(int), hipMemcpyDeviceToHost, s0.stream));
  if (parts)
    for (int i = 0; i < S; ++i)
      HIPCHK(hipMemcpyAsync(sres.data() + 8 * i, x.sp[i]->results.p, sizeof(double) * 8, hipMemcpyDeviceToHost,
                            s0.stream));
  HIPCHK(hipStreamSynchronize(s0.stream));
  if (bad) {
    HIPCHK(hipMemset(x.flag.p, 0, sizeof(int)));
    g_err = "Eta too large in FMT::dPHI";
    return CDFT_ETA_TOO_LARGE;
  }
  if (!parts)  // the fixed-mass multipliers need the sums
    for (int i = 0; i < S && with_dft; ++i)
      if (x.sp[i]->last_fixed_mass) {
        HIPCHK(hipMemcpy(sres.data() + 8 * i, x.sp[i]->results.p, sizeof(double) * 8, hipMemcpyDeviceToHost));
      }
  for (int i = 0; i < S; ++i)  // mu_ = sum dF rho / fixed mass (src/Species.cpp:99-110)
    if (x.sp[i]->last_fixed_mass) x.sp[i]->lagrange[0] = sres[8 * i + 7] / x.sp[i]->masses[0];
  if (parts) {
    const double F_hs = res * dV;
    double F_id = 0.0, F_ext = 0.0;  // species by species, as DFT::calculateFreeEnergyAndDerivatives adds them
    for (int i = 0; i < S; ++i) {
      const cdft_system_s& s = *x.sp[i];
      const double* r = sres.data() + 8 * i;
      if (with_dft && s.d.ideal && !no_ideal) F_id += r[1];
      if (with_dft) F_ext += csum({s.has_vext ? r[3] * dV : 0.0, -r[4] * dV * eval_mu(s)});
    }
    parts[0] = csum({F_id, F_hs, F_ext});
    parts[1] = F_id;
    parts[2] = F_hs;
    parts[3] = 0.0;
    parts[4] = F_ext;
  }
  return CDFT_OK;
}

}  // namespace cdft_impl

using namespace cdft_impl;

namespace {
// Adaptive Gauss-Kronrod (7/15) quadrature, absolute tolerance tol.
template <typename Fn>
double gk15(const Fn& f, double a, double b, double tol, int depth = 0) {
  static const double xk[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                               0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                               0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                               0.207784955007898467600689403773245, 0.0};
  static const double wk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                               0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                               0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                               0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
  static const double wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                               0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  double k = wk[7] * f(c), g = wg[3] * f(c);
  for (int i = 0; i < 7; ++i) {
    const double fp = f(c + h * xk[i]), fm = f(c - h * xk[i]);
    k += wk[i] * (fp + fm);
    if (i % 2 == 1) g += wg[i / 2] * (fp + fm);
  }
  k *= h;
  g *= h;
  if (std::fabs(k - g) <= tol || depth > 40) return k;
  return gk15(f, a, c, 0.5 * tol, depth + 1) + gk15(f, c, b, 0.5 * tol, depth + 1);
}
}  // namespace

extern "C" {

const char* cdft_last_error(void) { return g_err.c_str(); }

int cdft_device_count(int* n) {
  return guarded([&] {
    REQUIRE(n, "null argument");
    HIPCHK(hipGetDeviceCount(n));
    return CDFT_OK;
  });
}

int cdft_system_create(const cdft_system_desc* desc, cdft_system* out) {
  return guarded([&] {
    create_system(desc, out);
    return CDFT_OK;
  });
}

int cdft_system_destroy(cdft_system sys) {
  return guarded([&] {
    if (sys) {
      (void)hipStreamSynchronize(sys->stream);
      delete sys;
    }
    return CDFT_OK;
  });
}

int cdft_system_eval(cdft_system sys, const double* rho, double* dF, double* parts, int flags) {
  return guarded([&] { return run_eval(sys, rho, dF, parts, flags, true); });
}

int cdft_fmt_eval(cdft_system sys, const double* rho, double* dF, double* F_hs, int flags) {
  return guarded([&] {
    REQUIRE(sys && sys->m > 0, "system has no FMT functional");
    std::vector<double> parts(5 * (size_t)sys->nb);
    const int rc = run_eval(sys, rho, dF, parts.data(), flags, false);
    if (rc == CDFT_OK && F_hs)
      for (int b = 0; b < sys->nb; ++b) F_hs[b] = parts[5 * b + 2];
    return rc;
  });
}

namespace {
int run_hess(cdft_system s, const double* rho, const double* v, double* out, int flags, bool with_dft) {
  REQUIRE(s && rho && v && out, "null argument");
  const bool dev = flags & CDFT_DEVICE_PTRS;
  const bool acc = flags & CDFT_ACCUMULATE;
  const size_t bytes = sizeof(double) * s->N * s->nb;
  const double *rho_d = rho, *v_d = v;
  double* out_d = out;
  if (!dev) {
    if (s->vdir.bytes < bytes) s->vdir.alloc(bytes);
    HIPCHK(hipMemcpyAsync(s->rho.p, rho, bytes, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->vdir.p, v, bytes, hipMemcpyHostToDevice, s->stream));
    if (acc) HIPCHK(hipMemcpyAsync(s->dF.p, out, bytes, hipMemcpyHostToDevice, s->stream));
    rho_d = s->rho.as<double>();
    v_d = s->vdir.as<double>();
    out_d = s->dF.as<double>();
  }
  const bool fixed_bnd = with_dft && (s->d.boundary & CDFT_FIXED_BACKGROUND);
  if (fixed_bnd) {  // DFT::matrix_dot_v_intern with fixed boundaries, src/DFT.cpp:441-449
    const long nbnd = boundary_count(s->g.Nx, s->g.Ny, s->g.Nz);
    cons_sum(*s, v_d, nullptr, nbnd, CONS_BABS, 6);
    std::vector<double> res(8 * (size_t)s->nb);
    HIPCHK(hipMemcpyAsync(res.data(), s->results.p, sizeof(double) * 8 * s->nb, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int b = 0; b < s->nb; ++b)
      REQUIRE(res[8 * b + 6] == 0.0,
              "Input vector v must have zero boundary entries in DFT::hessian_dot_v when the species has fixed boundaries");
  }
  hess_device(*s, rho_d, v_d, out_d, with_dft, !(flags & CDFT_NO_IDEAL), acc);
  if (fixed_bnd) {  // src/DFT.cpp:478-485
    const long nbnd = boundary_count(s->g.Nx, s->g.Ny, s->g.Nz);
    hipLaunchKernelGGL(k_cons_bset, cons_grid(*s, nbnd), dim3(kThreads), 0, s->stream, out_d, nbnd, s->N, s->g.Nx,
                       s->g.Ny, s->g.Nz, nullptr, 8L, 0.0);
    HIPCHK(hipGetLastError());
  }
  if (!dev) HIPCHK(hipMemcpyAsync(out, s->dF.p, bytes, hipMemcpyDeviceToHost, s->stream));
  std::vector<double> parts(5 * (size_t)s->nb);
  const bool last = s->last_with_dft, last_fm = s->last_fixed_mass;
  s->last_fixed_mass = false;
  const int rc = fetch_parts(s, parts.data());  // waits; reports eta > 1
  s->last_with_dft = last;
  s->last_fixed_mass = last_fm;
  return rc;
}
}  // namespace

int cdft_system_hessian_dot(cdft_system sys, const double* rho, const double* v, double* out, int flags) {
  return guarded([&] { return run_hess(sys, rho, v, out, flags, true); });
}

// No counterpart in the reference: cdft_system_hessian_dot on device pointers for callers that cannot allocate.
// The spectra of the direction (A2: nb spectra, B2: nconv nb spectra of Nx Ny Zp complex doubles) live in this
// device's scratch instead of the system's lazily allocated buffers; every element the kernels read there is
// written earlier in the same call, so nothing needs clearing.  No hipMalloc / hipFree / host staging.
int cdft_system_hessian_dot_noalloc(cdft_system sys, const double* rho, const double* v, double* out, int flags) {
  return guarded([&] {
    REQUIRE(sys && rho && v && out, "null argument");
    cdft_system_s& s = *sys;
    REQUIRE(flags & CDFT_DEVICE_PTRS, "cdft_system_hessian_dot_noalloc takes device pointers (CDFT_DEVICE_PTRS)");
    REQUIRE(!(s.d.boundary & CDFT_FIXED_BACKGROUND),
            "cdft_system_hessian_dot_noalloc: fixed-background systems use cdft_system_hessian_dot");
    const size_t cs = sizeof(cplx<double>);
    const size_t need = cs * (size_t)s.Nkp * (size_t)s.nb * (1 + (size_t)s.nconv);
    if (need > kScratchBytes) {
      g_err = "cdft_system_hessian_dot_noalloc: the system needs " + std::to_string(need) +
              " bytes of work memory, more than the " + std::to_string(kScratchBytes) + " bytes of its scratch";
      return CDFT_BAD_ARG;
    }
    cplx<double>* A2 = static_cast<cplx<double>*>(scratch_of(s.device));
    cplx<double>* B2 = A2 + (size_t)s.Nkp * s.nb;
    hess_device(s, rho, v, out, true, !(flags & CDFT_NO_IDEAL), (flags & CDFT_ACCUMULATE) != 0, A2, B2);
    // eta > 1 anywhere: the flag the kernels raised (read without host containers)
    int fl[64];
    for (int b0 = 0; b0 < s.nb; b0 += 64) {
      const int n = std::min(64, s.nb - b0);
      HIPCHK(hipMemcpyAsync(fl, s.flag.as<int>() + b0, sizeof(int) * n, hipMemcpyDeviceToHost, s.stream));
      HIPCHK(hipStreamSynchronize(s.stream));
      for (int b = 0; b < n; ++b)
        if (fl[b]) {
          hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(kThreads), 0, s.stream, s.flag.as<int>(), s.nb);
          HIPCHK(hipStreamSynchronize(s.stream));
          g_err = "Eta too large in FMT::dPHI";
          return CDFT_ETA_TOO_LARGE;
        }
    }
    return CDFT_OK;
  });
}

int cdft_fmt_hessian_dot(cdft_system sys, const double* rho, const double* v, double* out, int flags) {
  return guarded([&] {
    REQUIRE(sys && sys->m > 0, "system has no FMT functional");
    return run_hess(sys, rho, v, out, flags, false);
  });
}

namespace {
// gx = g . x: DDFT::g_dot_x (src/DDFT.cpp:320-351, forward differences with the density averaged over each link,
// D = 1 / (2 dx^2) per axis, periodic neighbours via get_neighbors :294-312); zero on the boundary ix, iy or
// iz = 0 of a fixed background.
__global__ __launch_bounds__(256) void k_gdotx(const double* __restrict__ rho, const double* __restrict__ x,
                                               double* __restrict__ out, int Nx, int Ny, int Nz, double Dx, double Dy,
                                               double Dz, int fixed_bnd, int central) {
  const long N = (long)Nx * Ny * Nz;
  const int st = central ? 2 : 1;  // x's neighbours at the stride, the density's at one site (src/DDFT.cpp:325)
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (long)gridDim.x * blockDim.x) {
    const int z = p % Nz, y = (p / Nz) % Ny, ix = p / ((long)Nz * Ny);
    const long px = ((long)((ix + 1) % Nx) * Ny + y) * Nz + z, mx = ((long)((ix + Nx - 1) % Nx) * Ny + y) * Nz + z;
    const long py = ((long)ix * Ny + (y + 1) % Ny) * Nz + z, my = ((long)ix * Ny + (y + Ny - 1) % Ny) * Nz + z;
    const long pz = ((long)ix * Ny + y) * Nz + (z + 1) % Nz, mz = ((long)ix * Ny + y) * Nz + (z + Nz - 1) % Nz;
    const long qx = ((long)((ix + st) % Nx) * Ny + y) * Nz + z, nx = ((long)((ix + 2 * Nx - st) % Nx) * Ny + y) * Nz + z;
    const long qy = ((long)ix * Ny + (y + st) % Ny) * Nz + z, ny = ((long)ix * Ny + (y + 2 * Ny - st) % Ny) * Nz + z;
    const long qz = ((long)ix * Ny + y) * Nz + (z + st) % Nz, nz = ((long)ix * Ny + y) * Nz + (z + 2 * Nz - st) % Nz;
    const double x0 = x[p], d0 = central ? 0.0 : rho[p];
    double g = Dx * ((rho[px] + d0) * (x[qx] - x0) - (d0 + rho[mx]) * (x0 - x[nx])) +
               Dy * ((rho[py] + d0) * (x[qy] - x0) - (d0 + rho[my]) * (x0 - x[ny])) +
               Dz * ((rho[pz] + d0) * (x[qz] - x0) - (d0 + rho[mz]) * (x0 - x[nz]));
    if (fixed_bnd && (ix == 0 || y == 0 || z == 0)) g = 0.0;
    out[p] = g;
  }
}
__global__ __launch_bounds__(256) void k_negate(double* __restrict__ a, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = -a[i];
}

// r = A x as Dynamical_Matrix::matrix_dot_v1 applies it (src/Dynamical_Matrix.cpp:34-70): the input's boundary
// entries zeroed on a copy (fixed background), the product, the output's boundary entries zeroed.  A is the DFT
// Hessian (DFT::matrix_dot_v_intern, src/DFT.cpp:435-488; CDFT_NO_IDEAL = full_hessian_ false) or, with
// CDFT_DYNAMIC_MATRIX, the DDFT dynamical matrix g . (H x) (DDFT::matrix_dot_v_intern, src/DDFT.cpp:353-361);
// CDFT_SQUARED_MATRIX applies it twice, negated for the dynamical matrix (is_dynamic).  t: scratch of N values.
void matrix_apply(cdft_system_s& s, const double* rho, const double* x, double* r, double* t, int flags) {
  const long N = s.N;
  const size_t bytes = sizeof(double) * N;
  const bool bnd = s.d.boundary & CDFT_FIXED_BACKGROUND, ideal = !(flags & CDFT_NO_IDEAL);
  const bool dynamic = flags & CDFT_DYNAMIC_MATRIX, squared = flags & CDFT_SQUARED_MATRIX;
  const bool central = flags & CDFT_CENTRAL_DIFFERENCES;
  const long nbnd = boundary_count(s.g.Nx, s.g.Ny, s.g.Nz);
  const dim3 eg((unsigned)std::max<long>(1, std::min<long>(div_up(N, kThreads), 8192)));
  auto zero_bnd = [&](double* v) {
    if (!bnd) return;
    hipLaunchKernelGGL(k_cons_bset, cons_grid(s, nbnd), dim3(kThreads), 0, s.stream, v, nbnd, N, s.g.Nx, s.g.Ny,
                       s.g.Nz, nullptr, 8L, 0.0);
    HIPCHK(hipGetLastError());
  };
  auto once = [&]() {  // t -> r
    hess_device(s, rho, t, r, true, ideal, false);
    zero_bnd(r);
    if (dynamic) {
      HIPCHK(hipMemcpyAsync(t, r, bytes, hipMemcpyDeviceToDevice, s.stream));
      const double st = central ? 2.0 : 1.0;
      hipLaunchKernelGGL(k_gdotx, eg, dim3(kThreads), 0, s.stream, rho, t, r, s.g.Nx, s.g.Ny, s.g.Nz,
                         1.0 / (2 * st * s.d.dx * s.d.dx), 1.0 / (2 * st * s.d.dy * s.d.dy),
                         1.0 / (2 * st * s.d.dz * s.d.dz), bnd ? 1 : 0, central ? 1 : 0);
      HIPCHK(hipGetLastError());
    }
  };
  HIPCHK(hipMemcpyAsync(t, x, bytes, hipMemcpyDeviceToDevice, s.stream));
  zero_bnd(t);
  once();
  if (squared) {
    HIPCHK(hipMemcpyAsync(t, r, bytes, hipMemcpyDeviceToDevice, s.stream));
    once();
    if (dynamic) {
      hipLaunchKernelGGL(k_negate, eg, dim3(kThreads), 0, s.stream, r, N);
      HIPCHK(hipGetLastError());
    }
  }
  zero_bnd(r);
}
}  // namespace

namespace {
// One term of the Chebyshev recursion of Log_Det::calculate_log_det (src/Log_Det.cpp:122-153):
// x = (((rs r + z) scale) cA) + cB w1 - w0 (rs = -1 negates the product of a dynamical matrix) with z = zf * sum (the zero-mode shift lam_mid accu(w1) / Ntot, sum on
// the device; zf = 0 without it); w2 = x; u = c0 w1 + ck x on the first term (w1 = the probe, no w0), else
// u += ck x.
__global__ __launch_bounds__(256) void k_cheb_step(const double* __restrict__ r, const double* __restrict__ w1,
                                                   const double* __restrict__ w0, double* __restrict__ w2,
                                                   double* __restrict__ u, long n, const double* __restrict__ sum,
                                                   double zf, double scale, double cA, double cB, double c0,
                                                   double ck, int first, double rs) {
  const double z = zf != 0.0 ? zf * sum[0] : 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double x = rs * r[i] + z;
    x *= scale;
    x *= cA;
    x += cB * w1[i];
    if (w0) x -= w0[i];
    w2[i] = x;
    u[i] = first ? c0 * w1[i] + ck * x : u[i] + ck * x;
  }
}
}  // namespace

int cdft_system_log_det(cdft_system sys, const double* rho, int order, double lam_max, double lam_min, long seed,
                        int num_samples, int has_zero_eigenvalue, int flags, double* log_det, double* variance) {
  return guarded([&] {
    REQUIRE(sys && rho && log_det && variance, "null argument");
    cdft_system_s& s = *sys;
    REQUIRE(s.nb == 1, "the log-determinant takes one system, not a batch");
    REQUIRE(order >= 1 && num_samples >= 1, "order and num_samples must be positive");
    REQUIRE(lam_max > 0 && lam_min > 0 && lam_max > lam_min, "need 0 < lam_min < lam_max");
    const long N = s.N;
    const size_t bytes = sizeof(double) * N;
    const bool bnd = s.d.boundary & CDFT_FIXED_BACKGROUND;
    const long nbnd = boundary_count(s.g.Nx, s.g.Ny, s.g.Nz);
    // Log_Det::Log_Det and get_coefficients (src/Log_Det.cpp:22-62)
    const double scale = 1.0 / (lam_min + lam_max), a = lam_min * scale, b = lam_max * scale;
    const double lam_mid = (lam_max + lam_min) / 2;
    std::vector<double> c(order + 1, 0.0);
    for (int k = 0; k <= order; ++k) {
      const double xk = std::cos(M_PI * (k + 0.5) / (order + 1));
      const double f = std::log(0.5 * (b - a) * xk + 0.5 * (b + a));
      double tm2 = 1.0, tm1 = xk;
      for (int j = 0; j <= order; ++j) {
        const double t = j == 0 ? 1.0 : j == 1 ? xk : 2 * xk * tm1 - tm2;
        if (j >= 2) {
          tm2 = tm1;
          tm1 = t;
        }
        c[j] += f * t;
      }
    }
    for (int j = 0; j <= order; ++j) c[j] *= (j == 0 ? 1.0 : 2.0) / (order + 1);
    const double c1 = order >= 1 ? c[1] : 0.0;
    // device vectors: probe, three recursion terms, accumulator, product, projected input, sums
    DevBuf rho_buf, work, sums;
    work.alloc(7 * bytes);
    sums.alloc(2 * sizeof(double));
    double* W = work.as<double>();
    double *v = W, *w0 = W + N, *w1 = W + 2 * N, *w2 = W + 3 * N, *u = W + 4 * N, *r = W + 5 * N, *t = W + 6 * N;
    const double* rho_d = rho;
    if (!(flags & CDFT_DEVICE_PTRS)) {
      rho_buf.alloc(bytes);
      HIPCHK(hipMemcpyAsync(rho_buf.p, rho, bytes, hipMemcpyHostToDevice, s.stream));
      rho_d = rho_buf.as<double>();
    }
    const long nblk = std::min<long>(s.pslots, std::max<long>(1, std::min<long>(div_up(N, kThreads), 1024)));
    auto dsum = [&](const double* x, const double* y, int slot) {  // compensated sum (x) or dot (x, y)
      hipLaunchKernelGGL(k_cons_sum, dim3((unsigned)nblk, 1), dim3(kThreads), 0, s.stream, x, y, N, N,
                         y ? CONS_DOT : CONS_SUM, s.g.Nx, s.g.Ny, s.g.Nz, s.partials.as<double2>(), s.pslots);
      HIPCHK(hipGetLastError());
      hipLaunchKernelGGL(k_reduce_partials_wide, dim3(1), dim3(1024), 0, s.stream, s.partials.as<double2>(), nblk, 1,
                         sums.as<double>() + slot, s.pslots, 1L);
      HIPCHK(hipGetLastError());
    };
    auto zero_bnd = [&](double* x) {
      if (!bnd) return;
      hipLaunchKernelGGL(k_cons_bset, cons_grid(s, nbnd), dim3(kThreads), 0, s.stream, x, nbnd, N, s.g.Nx, s.g.Ny,
                         s.g.Nz, nullptr, 8L, 0.0);
      HIPCHK(hipGetLastError());
    };
    // Log_Det::matrix_dot_v1 (src/Log_Det.cpp:73-92): r = A x; a dynamical matrix's product is negated
    // (src/Log_Det.cpp:124, 139)
    auto matvec = [&](const double* x) { matrix_apply(s, rho_d, x, r, t, flags); };
    const double rs = (flags & CDFT_DYNAMIC_MATRIX) ? -1.0 : 1.0;
    const dim3 eg((unsigned)std::max<long>(1, std::min<long>(div_up(N, kThreads), 8192)));
    const double zf = has_zero_eigenvalue ? lam_mid / (double)N : 0.0;
    if (seed <= 0) {
      std::random_device rd;
      seed = rd();
    }
    std::mt19937 rng(seed);
    std::uniform_int_distribution<> distrib(0, 1);
    std::vector<double> vh(N);
    double ld = 0.0, var = 0.0, sd = 0.0;
    for (int i = 1; i <= num_samples; ++i) {
      for (long p = 0; p < N; ++p) vh[p] = distrib(rng) > 0 ? 1.0 : -1.0;
      HIPCHK(hipMemcpyAsync(v, vh.data(), bytes, hipMemcpyHostToDevice, s.stream));
      zero_bnd(v);
      matvec(v);
      if (zf != 0.0) dsum(v, nullptr, 0);
      hipLaunchKernelGGL(k_cheb_step, eg, dim3(kThreads), 0, s.stream, r, v, nullptr, w1, u, N, sums.as<double>(),
                         zf, scale, 2.0 / (b - a), -(b + a) / (b - a), c[0], c1, 1, rs);
      HIPCHK(hipGetLastError());
      double *p0 = v, *p1 = w1, *p2 = w2, *spare = w0;
      for (int k = 2; k <= order; ++k) {
        matvec(p1);
        if (zf != 0.0) dsum(p1, nullptr, 0);
        hipLaunchKernelGGL(k_cheb_step, eg, dim3(kThreads), 0, s.stream, r, p1, p0, p2, u, N, sums.as<double>(), zf,
                           scale, 4.0 / (b - a), -2 * (b + a) / (b - a), 0.0, c[k], 0, rs);
        HIPCHK(hipGetLastError());
        // w0 <- w1, w1 <- w2; the probe v stays intact for the final dot product
        double* freed = p0 == v ? spare : p0;
        p0 = p1;
        p1 = p2;
        p2 = freed;
      }
      dsum(v, u, 1);
      double vu = 0.0;
      HIPCHK(hipMemcpyAsync(&vu, sums.as<double>() + 1, sizeof(double), hipMemcpyDeviceToHost, s.stream));
      HIPCHK(hipStreamSynchronize(s.stream));
      ld += vu;
      var += vu * vu;
      const double av = ld / i, av2 = var / i;
      sd = std::sqrt(std::fabs(av2 - av * av));
    }
    double out = ld / num_samples;
    if (has_zero_eigenvalue) out -= std::log(lam_mid);
    out -= (double)(N - (bnd ? nbnd : 0)) * std::log(scale);
    const bool last = s.last_with_dft, last_fm = s.last_fixed_mass;
    s.last_fixed_mass = false;
    const int rc = fetch_parts(&s, nullptr);  // reports eta > 1 met by any product
    s.last_with_dft = last;
    s.last_fixed_mass = last_fm;
    if (rc != CDFT_OK) return rc;
    *log_det = out;
    *variance = sd;
    return (int)CDFT_OK;
  });
}

namespace {
// r <- (r + vs (0.5 sum[0])) * sgn_scale   (src/Eigenvalues.cpp:49-53; vs null: no shift)
__global__ __launch_bounds__(256) void k_eig_shift_scale(double* __restrict__ r, const double* __restrict__ vs,
                                                         const double* __restrict__ sum, double sgn_scale, long n) {
  const double h = vs ? 0.5 * sum[0] : 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double x = r[i];
    if (vs) x += h * vs[i];
    r[i] = x * sgn_scale;
  }
}
// df = (r - f x) (2 / x2), f = sum[1] / x2, x2 = sum[2]   (src/Eigenvalues.cpp:55-59)
__global__ __launch_bounds__(256) void k_eig_grad(const double* __restrict__ r, const double* __restrict__ x,
                                                  const double* __restrict__ sum, double* __restrict__ df, long n) {
  const double x2 = sum[2], f = sum[1] / x2, g = 2 / x2;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    df[i] = (r[i] + (-f) * x[i]) * g;
}
}  // namespace

int cdft_system_dynamic_dot(cdft_system sys, const double* rho, const double* v, double* out, int flags) {
  return guarded([&] {
    REQUIRE(sys && rho && v && out, "null argument");
    cdft_system_s& s = *sys;
    REQUIRE(s.nb == 1, "the dynamical matrix takes one system, not a batch");
    const long N = s.N;
    const size_t bytes = sizeof(double) * N;
    const bool dev = flags & CDFT_DEVICE_PTRS;
    DevBuf work;
    work.alloc((dev ? 1 : 4) * bytes);
    double* t = work.as<double>();
    const double *rho_d = rho, *v_d = v;
    double* out_d = out;
    if (!dev) {
      double* h = t + N;
      HIPCHK(hipMemcpyAsync(h, rho, bytes, hipMemcpyHostToDevice, s.stream));
      HIPCHK(hipMemcpyAsync(h + N, v, bytes, hipMemcpyHostToDevice, s.stream));
      rho_d = h;
      v_d = h + N;
      out_d = h + 2 * N;
    }
    matrix_apply(s, rho_d, v_d, out_d, t,
                 (flags & (CDFT_NO_IDEAL | CDFT_SQUARED_MATRIX | CDFT_CENTRAL_DIFFERENCES)) | CDFT_DYNAMIC_MATRIX);
    if (!dev) HIPCHK(hipMemcpyAsync(out, out_d, bytes, hipMemcpyDeviceToHost, s.stream));
    const bool last = s.last_with_dft, last_fm = s.last_fixed_mass;
    s.last_fixed_mass = false;
    const int rc = fetch_parts(&s, nullptr);  // waits; reports eta > 1
    s.last_with_dft = last;
    s.last_fixed_mass = last_fm;
    return rc;
  });
}

int cdft_system_eigen_objective(cdft_system sys, const double* rho, const double* x, const double* vshift,
                                double scale, int change_sign, int flags, double* f, double* df) {
  return guarded([&] {
    REQUIRE(sys && rho && x && f && df, "null argument");
    cdft_system_s& s = *sys;
    REQUIRE(s.nb == 1, "the eigenvalue objective takes one system, not a batch");
    const long N = s.N;
    const size_t bytes = sizeof(double) * N;
    const bool dev = flags & CDFT_DEVICE_PTRS;
    DevBuf work, sums, in;
    work.alloc(2 * bytes + (dev ? 0 : 3 * bytes));
    sums.alloc(3 * sizeof(double));
    double *r = work.as<double>(), *t = r + N;
    const double *rho_d = rho, *x_d = x, *vs_d = vshift;
    double* df_d = df;
    if (!dev) {
      double* h = t + N;
      HIPCHK(hipMemcpyAsync(h, rho, bytes, hipMemcpyHostToDevice, s.stream));
      HIPCHK(hipMemcpyAsync(h + N, x, bytes, hipMemcpyHostToDevice, s.stream));
      rho_d = h;
      x_d = h + N;
      df_d = h + 2 * N;
      if (vshift) {
        in.alloc(bytes);
        HIPCHK(hipMemcpyAsync(in.p, vshift, bytes, hipMemcpyHostToDevice, s.stream));
        vs_d = in.as<double>();
      }
    }
    const long nblk = std::min<long>(s.pslots, std::max<long>(1, std::min<long>(div_up(N, kThreads), 1024)));
    auto ddot = [&](const double* a, const double* b, int slot) {
      hipLaunchKernelGGL(k_cons_sum, dim3((unsigned)nblk, 1), dim3(kThreads), 0, s.stream, a, b, N, N, CONS_DOT,
                         s.g.Nx, s.g.Ny, s.g.Nz, s.partials.as<double2>(), s.pslots);
      HIPCHK(hipGetLastError());
      hipLaunchKernelGGL(k_reduce_partials_wide, dim3(1), dim3(1024), 0, s.stream, s.partials.as<double2>(), nblk, 1,
                         sums.as<double>() + slot, s.pslots, 1L);
      HIPCHK(hipGetLastError());
    };
    matrix_apply(s, rho_d, x_d, r, t, flags);  // Dynamical_Matrix::matrix_dot_v1: r = A x
    if (vs_d) ddot(vs_d, x_d, 0);
    const dim3 eg((unsigned)std::max<long>(1, std::min<long>(div_up(N, kThreads), 8192)));
    hipLaunchKernelGGL(k_eig_shift_scale, eg, dim3(kThreads), 0, s.stream, r, vs_d, sums.as<double>(),
                       (change_sign ? -1.0 : 1.0) * scale, N);
    HIPCHK(hipGetLastError());
    ddot(x_d, r, 1);
    ddot(x_d, x_d, 2);
    hipLaunchKernelGGL(k_eig_grad, eg, dim3(kThreads), 0, s.stream, r, x_d, sums.as<double>(), df_d, N);
    HIPCHK(hipGetLastError());
    double sh[3];
    HIPCHK(hipMemcpyAsync(sh, sums.p, sizeof(sh), hipMemcpyDeviceToHost, s.stream));
    if (!dev) HIPCHK(hipMemcpyAsync(df, df_d, bytes, hipMemcpyDeviceToHost, s.stream));
    const bool last = s.last_with_dft, last_fm = s.last_fixed_mass;
    s.last_fixed_mass = false;
    const int rc = fetch_parts(&s, nullptr);  // waits; reports eta > 1
    s.last_with_dft = last;
    s.last_fixed_mass = last_fm;
    if (rc != CDFT_OK) return rc;
    *f = sh[1] / sh[2];
    return (int)CDFT_OK;
  });
}

int cdft_system_eval_async(cdft_system sys, const double* rho_d, double* dF_d, int flags) {
  return guarded([&] {
    REQUIRE(sys && rho_d && dF_d, "null argument");
    const bool fmt_only = flags & CDFT_FMT_ONLY;
    REQUIRE(!fmt_only || sys->m > 0, "system has no FMT functional");
    sys->no_ideal = !fmt_only && (flags & CDFT_NO_IDEAL);
    enqueue_eval(sys, rho_d, dF_d, !fmt_only, (flags & CDFT_ACCUMULATE) != 0);
    sys->no_ideal = false;
    return CDFT_OK;
  });
}

int cdft_system_fetch(cdft_system sys, double* parts) {
  return guarded([&] {
    REQUIRE(sys, "null argument");
    return fetch_parts(sys, parts);
  });
}

int cdft_system_time(cdft_system sys, const double* rho_d, double* dF_d, int flags, int iters, double* ms) {
  return guarded([&] {
    REQUIRE(sys && rho_d && dF_d && ms && iters > 0, "bad argument");
    const bool fmt_only = flags & CDFT_FMT_ONLY;
    REQUIRE(!fmt_only || sys->m > 0, "system has no FMT functional");
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, sys->stream));
    for (int i = 0; i < iters; ++i) enqueue_eval(sys, rho_d, dF_d, !fmt_only, (flags & CDFT_ACCUMULATE) != 0);
    HIPCHK(hipEventRecord(e1, sys->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t;
    return CDFT_OK;
  });
}

int cdft_malloc(void** p, unsigned long long bytes) {
  return guarded([&] {
    REQUIRE(p, "null argument");
    HIPCHK(hipMalloc(p, bytes));
    return CDFT_OK;
  });
}

int cdft_free(void* p) {
  return guarded([&] {
    if (p) HIPCHK(hipFree(p));
    return CDFT_OK;
  });
}

int cdft_memcpy_h2d(void* dst, const void* src, unsigned long long bytes) {
  return guarded([&] {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return CDFT_OK;
  });
}

int cdft_memcpy_d2h(void* dst, const void* src, unsigned long long bytes) {
  return guarded([&] {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return CDFT_OK;
  });
}


int cdft_potential_hsd(const cdft_potential_desc* p, double kT, double* hsd) {
  return guarded([&] {
    REQUIRE(p && hsd && p->sigma > 0 && kT > 0, "bad arguments");
    const Pot v = pot_from(*p);
    const double hc = v.hard_core(), rlimit = v.bh ? v.r0 : v.rmin;
    auto f = [&](double r) { return r <= 0 ? 1.0 : 1.0 - std::exp(-v.V0(r) / kT); };
    *hsd = hc + (rlimit > hc ? gk15(f, hc, rlimit, 1e-13) : 0.0);
    return (int)CDFT_OK;
  });
}

int cdft_potential_vdw(const cdft_potential_desc* p, double kT, double* a) {
  return guarded([&] {
    REQUIRE(p && a && p->sigma > 0 && kT > 0, "bad arguments");
    const Pot v = pot_from(*p);
    auto f = [&](double r) { return r * r * v.watt2(r * r); };
    double s = 0.0;
    if (v.bh) s = gk15(f, v.r0, v.rcut, 1e-14);
    else s = gk15(f, v.r_att_min, v.rmin, 1e-14) + gk15(f, v.rmin, v.rcut, 1e-14);
    *a = (2 * M_PI / kT) * s;
    return (int)CDFT_OK;
  });
}

int cdft_potential_watt(const cdft_potential_desc* p, int n, const double* r, double* w) {
  return guarded([&] {
    REQUIRE(p && r && w && n >= 0, "bad arguments");
    const Pot v = pot_from(*p);
    for (int i = 0; i < n; ++i) w[i] = v.watt2(r[i] * r[i]);
    return (int)CDFT_OK;
  });
}

int cdft_lj_hsd(double sigma, double eps, double rcut, double kT, double* hsd) {
  const cdft_potential_desc p{CDFT_POT_LJ, sigma, eps, rcut, 0.0, 0, 0.0};
  return cdft_potential_hsd(&p, kT, hsd);
}

int cdft_system_bulk(cdft_system sys, double rho, double* f, double* mu) {
  return guarded([&] {
    REQUIRE(sys && rho > 0, "bad argument");
    // uniform weighted densities n_a = rho * sum_R w_a(R) (fillUniform,
    // include/Fundamental_Measures.h:20-31); derivatives collapse onto them.
    const double d = sys->d.hsd;
    double fex = 0.0, muex = 0.0, fao = 0.0, muao = 0.0;
    auto omega = [](double dd, double* w) {
      const double v[11] = {M_PI * dd * dd * dd / 6, M_PI * dd * dd, 0, 0, 0, M_PI * dd * dd / 3, 0, 0,
                            M_PI * dd * dd / 3, 0, M_PI * dd * dd / 3};
      for (int a = 0; a < 11; ++a) w[a] = v[a];
    };
    if (sys->m > 0) {
      double om[11];
      omega(d, om);
      double n[11], out[11];
      for (int a = 0; a < 11; ++a) n[a] = rho * om[a];
      const PhiParams pp = make_phi_params(d, sys->d.A, sys->d.B,
                                           sys->d.model == CDFT_WHITEBEAR_I || sys->d.model == CDFT_WHITEBEAR_II);
      switch (sys->d.model) {
        case CDFT_ROSENFELD: fex = phi_site<M_ROSENFELD>(n, pp, out); break;
        case CDFT_RSLT: fex = phi_site<M_RSLT>(n, pp, out); break;
        case CDFT_ESFMT: fex = phi_site<M_ESFMT>(n, pp, out); break;
        case CDFT_WHITEBEAR_I: fex = phi_site<M_WBI>(n, pp, out); break;
        default: fex = phi_site<M_WBII>(n, pp, out); break;
      }
      for (int a = 0; a < 11; ++a) muex += out[a] * om[a];
      if (sys->has_ao) {
        const double dp = 2 * sys->d.ao_Rp, rp = sys->d.ao_reservoir;
        double omp[11], arg = 0.0, pref = 0.0;
        omega(dp, omp);
        switch (sys->d.model) {
          case CDFT_ROSENFELD: bulk_ao<M_ROSENFELD>(om, omp, rho, pp, dp, &arg, &pref); break;
          case CDFT_RSLT: bulk_ao<M_RSLT>(om, omp, rho, pp, dp, &arg, &pref); break;
          case CDFT_ESFMT: bulk_ao<M_ESFMT>(om, omp, rho, pp, dp, &arg, &pref); break;
          case CDFT_WHITEBEAR_I: bulk_ao<M_WBI>(om, omp, rho, pp, dp, &arg, &pref); break;
          default: bulk_ao<M_WBII>(om, omp, rho, pp, dp, &arg, &pref); break;
        }
        fao = rp * -std::expm1(-arg);
        muao = rp * pref * std::exp(-arg);
      }
    }
    const double a = sys->d.mean_field ? sys->a_vdw : 0.0;
    if (f) *f = csum({rho * std::log(rho) - rho, fex, fao, 0.5 * a * rho * rho});
    if (mu) *mu = csum({std::log(rho), muex, muao, a * rho});
    return CDFT_OK;
  });
}

int cdft_mf_kernel_points(int Nx, int Ny, int Nz, double dx, double dy, double dz, const cdft_potential_desc* pot,
                          const double* points, int npoints, int nradial, double* w_out) {
  return guarded([&] {
    REQUIRE(pot && points && w_out && npoints > 0 && Nx > 0 && Ny > 0 && Nz > 0, "bad argument");
    REQUIRE(dx > 0 && dy > 0 && dz > 0, "lattice spacings must be positive");
    const Pot v = pot_from(*pot);
    const int nr = nradial > 0 ? nradial : 512;  // Nr = 128*4, src/Interaction.cpp:769
    std::vector<double> x(nr), w(nr);
    gauss_legendre01(nr, x.data(), w.data());
    for (int k = 0; k < nr; ++k) {  // gsl_integration_glfixed_point(0, Rc, k, ...)
      x[k] *= v.rcut;
      w[k] *= v.rcut;
    }
    const long N = (long)Nx * Ny * Nz;
    OwnedStream owned;
    hipStream_t st = owned.s;
    DevBuf dp, dr, dw, out;
    dp.alloc(sizeof(double) * 4 * (size_t)npoints);
    dr.alloc(sizeof(double) * nr);
    dw.alloc(sizeof(double) * nr);
    out.alloc(sizeof(double) * N);
    HIPCHK(hipMemcpyAsync(dp.p, points, sizeof(double) * 4 * (size_t)npoints, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dr.p, x.data(), sizeof(double) * nr, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dw.p, w.data(), sizeof(double) * nr, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(out.p, 0, sizeof(double) * N, st));
    const long total = (long)npoints * 48 * nr;
    hipLaunchKernelGGL(k_mf_points, dim3((unsigned)std::min<long>(div_up(total, 256), 65536)), dim3(256), 0, st,
                       dp.as<double>(), (long)npoints, dr.as<double>(), dw.as<double>(), nr, v, Nx, Ny, Nz, dx, dy, dz,
                       out.as<double>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(w_out, out.p, sizeof(double) * N, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return CDFT_OK;
  });
}

int cdft_system_set_fixed_masses(cdft_system sys, const double* masses) {
  return guarded([&] {
    REQUIRE(sys && masses, "null argument");
    REQUIRE(sys->d.fixed_mass > 0, "the system was created without a fixed mass");
    for (int b = 0; b < sys->nb; ++b) REQUIRE(masses[b] > 0, "fixed masses must be positive");
    HIPCHK(hipStreamSynchronize(sys->stream));
    sys->masses.assign(masses, masses + sys->nb);
    HIPCHK(hipMemcpy(sys->masses_d.p, masses, sizeof(double) * sys->nb, hipMemcpyHostToDevice));
    return CDFT_OK;
  });
}

int cdft_system_chem_potential(cdft_system sys, double* mu) {
  return guarded([&] {
    REQUIRE(sys && mu, "null argument");
    for (int b = 0; b < sys->nb; ++b) mu[b] = sys->lagrange[b];
    return CDFT_OK;
  });
}

int cdft_system_stream(cdft_system sys, void** hip_stream) {
  return guarded([&] {
    REQUIRE(sys && hip_stream, "null argument");
    *hip_stream = (void*)sys->stream;
    return CDFT_OK;
  });
}

int cdft_system_set_stream(cdft_system sys, void* hip_stream) {
  return guarded([&] {
    REQUIRE(sys, "null argument");
    REQUIRE(sys->own_stream, "the members of a mixture run on the mixture's stream");
    HIPCHK(hipStreamSynchronize(sys->stream));  // the work already enqueued ends before the switch
    sys->stream = hip_stream ? (hipStream_t)hip_stream : sys->own_stream;
    return CDFT_OK;
  });
}

int cdft_set_device(int device) {
  return guarded([&] {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    REQUIRE(device >= 0 && device < n, "no such device");
    HIPCHK(hipSetDevice(device));
    return CDFT_OK;
  });
}

int cdft_get_device(int* device) {
  return guarded([&] {
    REQUIRE(device, "null argument");
    HIPCHK(hipGetDevice(device));
    return CDFT_OK;
  });
}

int cdft_system_layout(cdft_system sys, int* stored, int* convolved) {
  return guarded([&] {
    REQUIRE(sys && stored && convolved, "null argument");
    *stored = sys->m;
    *convolved = sys->nconv;
    return (int)CDFT_OK;
  });
}

int cdft_mixture_create(const cdft_mixture_desc* desc, cdft_mixture* out) {
  return guarded([&] {
    create_mixture(desc, out);
    return (int)CDFT_OK;
  });
}

int cdft_mixture_destroy(cdft_mixture mix) {
  return guarded([&] {
    if (mix) {
      (void)hipStreamSynchronize(mix->sp[0]->stream);
      delete mix;
    }
    return (int)CDFT_OK;
  });
}

int cdft_mixture_chem_potential(cdft_mixture mix, double* mu) {
  return guarded([&] {
    REQUIRE(mix && mu, "null argument");
    for (size_t i = 0; i < mix->sp.size(); ++i) {
      const cdft_system_s& s = *mix->sp[i];
      mu[i] = s.last_fixed_mass ? s.lagrange[0] : s.d.mu;
    }
    return (int)CDFT_OK;
  });
}

int cdft_mixture_species(cdft_mixture mix, int* nspecies) {
  return guarded([&] {
    REQUIRE(mix && nspecies, "null argument");
    *nspecies = (int)mix->sp.size();
    return (int)CDFT_OK;
  });
}

int cdft_mixture_fmt_eval(cdft_mixture mix, const double* const* rho, double* const* dF, double* F_hs, int flags) {
  return guarded([&] {
    REQUIRE(mix && rho && dF, "null argument");
    const int S = (int)mix->sp.size();
    for (int i = 0; i < S; ++i) REQUIRE(rho[i] && dF[i], "null species array");
    const bool acc = flags & CDFT_ACCUMULATE;
    double parts[5];
    if (flags & CDFT_DEVICE_PTRS) {
      const int rc = mixture_eval(*mix, rho, dF, acc, parts);
      if (rc == CDFT_OK && F_hs) *F_hs = parts[2];
      return rc;
    }
    // host data: staged in each member's density and force buffers
    cdft_system_s& s0 = *mix->sp[0];
    const size_t bytes = sizeof(double) * s0.N;
    std::vector<const double*> rd(S);
    std::vector<double*> fd(S);
    for (int i = 0; i < S; ++i) {
      cdft_system_s& s = *mix->sp[i];
      HIPCHK(hipMemcpyAsync(s.rho.p, rho[i], bytes, hipMemcpyHostToDevice, s0.stream));
      if (acc) HIPCHK(hipMemcpyAsync(s.dF.p, dF[i], bytes, hipMemcpyHostToDevice, s0.stream));
      rd[i] = s.rho.as<double>();
      fd[i] = s.dF.as<double>();
    }
    const int rc = mixture_eval(*mix, rd.data(), fd.data(), acc, parts);
    if (rc != CDFT_OK) return rc;
    if (F_hs) *F_hs = parts[2];
    for (int i = 0; i < S; ++i)
      HIPCHK(hipMemcpyAsync(dF[i], mix->sp[i]->dF.p, bytes, hipMemcpyDeviceToHost, s0.stream));
    HIPCHK(hipStreamSynchronize(s0.stream));
    return (int)CDFT_OK;
  });
}

int cdft_mixture_eval(cdft_mixture mix, const double* const* rho, double* const* dF, double* parts, int flags) {
  return guarded([&] {
    REQUIRE(mix && rho && dF, "null argument");
    const int S = (int)mix->sp.size();
    for (int i = 0; i < S; ++i) REQUIRE(rho[i] && dF[i], "null species array");
    const bool acc = flags & CDFT_ACCUMULATE, no_ideal = flags & CDFT_NO_IDEAL;
    if (flags & CDFT_DEVICE_PTRS) return mixture_eval(*mix, rho, dF, acc, parts, true, no_ideal);
    cdft_system_s& s0 = *mix->sp[0];
    const size_t bytes = sizeof(double) * s0.N;
    std::vector<const double*> rd(S);
    std::vector<double*> fd(S);
    for (int i = 0; i < S; ++i) {
      cdft_system_s& s = *mix->sp[i];
      HIPCHK(hipMemcpyAsync(s.rho.p, rho[i], bytes, hipMemcpyHostToDevice, s0.stream));
      if (acc) HIPCHK(hipMemcpyAsync(s.dF.p, dF[i], bytes, hipMemcpyHostToDevice, s0.stream));
      rd[i] = s.rho.as<double>();
      fd[i] = s.dF.as<double>();
    }
    const int rc = mixture_eval(*mix, rd.data(), fd.data(), acc, parts, true, no_ideal);
    if (rc != CDFT_OK) return rc;
    for (int i = 0; i < S; ++i)
      HIPCHK(hipMemcpyAsync(dF[i], mix->sp[i]->dF.p, bytes, hipMemcpyDeviceToHost, s0.stream));
    HIPCHK(hipStreamSynchronize(s0.stream));
    return (int)CDFT_OK;
  });
}

int cdft_system_a_vdw(cdft_system sys, double* a) {
  return guarded([&] {
    REQUIRE(sys && a, "null argument");
    *a = sys->a_vdw;
    return CDFT_OK;
  });
}

int cdft_system_get_weights(cdft_system sys, double* w11) {
  return guarded([&] {
    REQUIRE(sys && w11, "null argument");
    REQUIRE(sys->m > 0, "system has no FMT functional");
    DevBuf wreal;
    wreal.alloc(sizeof(double) * 11 * sys->N);
    cell_weights(*sys, sys->d.hsd, wreal.as<double>());
    HIPCHK(hipMemcpyAsync(w11, wreal.p, sizeof(double) * 11 * sys->N, hipMemcpyDeviceToHost, sys->stream));
    HIPCHK(hipStreamSynchronize(sys->stream));
    return CDFT_OK;
  });
}

int cdft_system_get_mf_kernel(cdft_system sys, double* w) {
  return guarded([&] {
    REQUIRE(sys && w, "null argument");
    REQUIRE(sys->d.mean_field, "system has no mean field");
    DevBuf buf;
    buf.alloc(sizeof(double) * sys->N);
    launch_mf_kernel(*sys, buf.as<double>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(w, buf.p, sizeof(double) * sys->N, hipMemcpyDeviceToHost, sys->stream));
    HIPCHK(hipStreamSynchronize(sys->stream));
    const double f = mf_scale(sys->d);
    if (f != 1.0)
      for (long i = 0; i < sys->N; ++i) w[i] *= f;
    return CDFT_OK;
  });
}

int cdft_fft_r2c(int Nx, int Ny, int Nz, int precision, const double* in, double* out) {
  return guarded([&] {
    REQUIRE(in && out, "null argument");
    return standalone_fft(Nx, Ny, Nz, precision, in, out, true);
  });
}

int cdft_fft_c2r(int Nx, int Ny, int Nz, int precision, const double* in, double* out) {
  return guarded([&] {
    REQUIRE(in && out, "null argument");
    return standalone_fft(Nx, Ny, Nz, precision, in, out, false);
  });
}

int cdft_system_set_timing(cdft_system sys, int enable) {
  return guarded([&] {
    REQUIRE(sys, "null argument");
    sys->timing = enable != 0;
    return CDFT_OK;
  });
}

int cdft_system_get_timing(cdft_system sys, int* n, double* ms, const char** names) {
  return guarded([&] {
    REQUIRE(sys && n, "null argument");
    HIPCHK(hipStreamSynchronize(sys->stream));
    // segments between consecutive marks, summed per stage name in the order the names first appear
    std::vector<const char*> nm;
    std::vector<double> tot;
    for (int i = 0; i + 1 < sys->n_marks; ++i) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, sys->ev[i], sys->ev[i + 1]));
      const char* name = sys->ev_names[i + 1];
      size_t j = 0;
      while (j < nm.size() && std::strcmp(nm[j], name) != 0) ++j;
      if (j == nm.size()) {
        nm.push_back(name);
        tot.push_back(0.0);
      }
      tot[j] += t;
    }
    const int k = (int)std::min<size_t>(nm.size(), 16);
    for (int i = 0; i < k; ++i) {
      if (ms) ms[i] = tot[i];
      if (names) names[i] = nm[i];
    }
    *n = k;
    return CDFT_OK;
  });
}

}  // extern "C"
