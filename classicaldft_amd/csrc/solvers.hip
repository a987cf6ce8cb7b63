// FIRE2 minimizer and DDFT integrating-factor stepper on the device.
//
// The control logic (time-step adaptation, back-tracking, fixed-point
// iteration) runs on the host exactly as in the reference; every N-length
// vector operation, reduction and transform runs in HIP kernels on the
// system's stream, and the F + dF evaluations go through the fused pipeline.
//
// FIRE2: fireMinimizer2 (include/Minimizer.h:86-182, src/Minimizer.cpp:117-327)
// on the alias variables of FMT_Species (src/FMT_Species.cpp:29-93).
// DDFT: DDFT::step and helpers (src/DDFT.cpp:42-351), single species.
#include "system.hpp"

namespace cdft_impl {
namespace {

constexpr int kRedBlocks = 1024;

// Block-reduce NS compensated sums and NM maxima; slot layout per block:
// NS (s, c) pairs then NM (max, 0) pairs.
template <int NS, int NM>
__device__ void reduce_store(CSum* s, double* mx, double2* __restrict__ partials) {
  __shared__ double rs[NS > 0 ? NS : 1][kThreads], rc[NS > 0 ? NS : 1][kThreads], rm[NM > 0 ? NM : 1][kThreads];
  const int t = threadIdx.x;
  for (int q = 0; q < NS; ++q) { rs[q][t] = s[q].s; rc[q][t] = s[q].c; }
  for (int q = 0; q < NM; ++q) rm[q][t] = mx[q];
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (t < off) {
      for (int q = 0; q < NS; ++q) {
        CSum a{rs[q][t], rc[q][t]};
        a.merge(rs[q][t + off], rc[q][t + off]);
        rs[q][t] = a.s;
        rc[q][t] = a.c;
      }
      for (int q = 0; q < NM; ++q) rm[q][t] = fmax(rm[q][t], rm[q][t + off]);
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int q = 0; q < NS; ++q) partials[blockIdx.x * (NS + NM) + q] = make_double2(rs[q][0], rc[q][0]);
    for (int q = 0; q < NM; ++q) partials[blockIdx.x * (NS + NM) + NS + q] = make_double2(rm[q][0], 0.0);
  }
}

// out[q] for q < NS: compensated sum over blocks; q >= NS: max over blocks.
__global__ void k_finish(const double2* __restrict__ partials, int nblk, int NS, int NM, double* __restrict__ out) {
  __shared__ double rs[kThreads], rc[kThreads];
  const int t = threadIdx.x;
  const int nq = NS + NM;
  for (int q = 0; q < nq; ++q) {
    CSum a{0.0, 0.0};
    double m = 0.0;
    for (int b = t; b < nblk; b += kThreads) {
      const double2 p = partials[b * nq + q];
      if (q < NS) a.merge(p.x, p.y);
      else m = fmax(m, p.x);
    }
    rs[t] = q < NS ? a.s : m;
    rc[t] = a.c;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
      if (t < off) {
        if (q < NS) {
          CSum b{rs[t], rc[t]};
          b.merge(rs[t + off], rc[t + off]);
          rs[t] = b.s;
          rc[t] = b.c;
        } else {
          rs[t] = fmax(rs[t], rs[t + off]);
        }
      }
      __syncthreads();
    }
    if (t == 0) out[q] = q < NS ? rs[0] + rc[0] : rs[0];
    __syncthreads();
  }
}

__global__ void k_zero_flag(int* p) { *p = 0; }

// ----------------------------- FIRE2 kernels -------------------------------
// The alias maps of the species whose variables a FIRE run concatenates (one species: one segment): element i
// belongs to species i / nseg, each with its own c (FMT_Species, src/FMT_Species.cpp:29-93).
constexpr int kMaxAliasSeg = 8;
struct AliasSet {
  Alias a[kMaxAliasSeg];
  long nseg;  // elements per species
  int ns;     // species
  // a chain of selects on constant indices: no division and no dynamic indexing of the argument block
  __device__ __forceinline__ Alias of(long i) const {
    Alias r = a[0];
#pragma unroll
    for (int s = 1; s < kMaxAliasSeg; ++s)
      if (s < ns && i >= s * nseg) r = a[s];
    return r;
  }
};

__global__ void k_alias_from_rho(const double* __restrict__ rho, double* __restrict__ x, long N, AliasSet al) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads)
    x[i] = al.of(i).x_of(rho[i]);
}

__global__ void k_rho_from_alias(const double* __restrict__ x, double* __restrict__ rho, long N, AliasSet al) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads)
    rho[i] = al.of(i).rho(x[i]);
}

// dF -> alias derivative (in place) and the reductions that follow a
// getDF_DX: sum dF_a^2, sum v.dF_a, max|dF_a|, max|dF_raw|.
__global__ void k_fire_post(const double* __restrict__ x, const double* __restrict__ v, const double* __restrict__ raw,
                            double* __restrict__ dFa, long N, AliasSet al, double2* __restrict__ partials) {
  CSum s[2] = {{0, 0}, {0, 0}};
  double mx[2] = {0.0, 0.0};
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) {
    const double r = raw[i];
    const double a = r * al.of(i).dfac(x[i]);
    dFa[i] = a;
    s[0].add(a * a);
    s[1].add(v[i] * a);
    mx[0] = fmax(mx[0], fabs(a));
    mx[1] = fmax(mx[1], fabs(r));
  }
  reduce_store<2, 2>(s, mx, partials);
}

// v <- v - dt dF, with sum v^2 and sum dF^2 (SemiImplicitEuler, Minimizer.cpp:277-292).
__global__ void k_fire_kick(double* __restrict__ v, const double* __restrict__ dF, long N, double dt,
                            double2* __restrict__ partials) {
  CSum s[2] = {{0, 0}, {0, 0}};
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) {
    const double f = dF[i];
    const double nv = v[i] + f * (-dt);
    v[i] = nv;
    s[0].add(nv * nv);
    s[1].add(f * f);
  }
  reduce_store<2, 0>(s, nullptr, partials);
}

// v <- (1-a) v + mix dF; x <- x + dt v; rho <- alias(x) (Minimizer.cpp:295-300, FMT_Species.cpp:29-48).
__global__ void k_fire_mix(double* __restrict__ x, double* __restrict__ v, const double* __restrict__ dF,
                           double* __restrict__ rho, long N, double alpha, double mix, double dt, AliasSet al) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) {
    double nv = v[i] * (1 - alpha);
    nv += dF[i] * mix;
    v[i] = nv;
    const double nx = x[i] + nv * dt;
    x[i] = nx;
    rho[i] = al.of(i).rho(nx);
  }
}

// uphill: x <- x - dt/2 v; v <- 0.1 v (Minimizer.cpp:218-222)
__global__ void k_fire_uphill(double* __restrict__ x, double* __restrict__ v, long N, double dt) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) {
    x[i] += v[i] * (-0.5 * dt);
    v[i] *= 0.1;
  }
}

__global__ void k_copy4(const double* __restrict__ a, const double* __restrict__ b, const double* __restrict__ c,
                        const double* __restrict__ d, double* __restrict__ a2, double* __restrict__ b2,
                        double* __restrict__ c2, double* __restrict__ d2, long N, double bscale) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) {
    a2[i] = a[i];
    b2[i] = b[i] * bscale;
    c2[i] = c[i];
    d2[i] = d[i];
  }
}

// ----------------------------- DDFT kernels --------------------------------
// r = (g . dF)/dV - Lap_h rho, with max |(g . dF)/dV|:  g_dot_x (DDFT.cpp:320-351; zero on the boundary
// ix, iy or iz = 0 of a fixed background, :331-337), calculate_excess_RHS (:212-220), subtract_ideal_gas (:223-235).
// With central differences (DDFT(dft, showGraphics, true), :325, 345) dF's neighbours are two sites away, the
// link densities lose the site's own rho and D = 1/(4 dx^2) (passed in Dx, Dy, Dz); the density's neighbours stay
// one site away, as in the ideal-gas Laplacian.
__global__ void k_ddft_rhs(const double* __restrict__ rho, const double* __restrict__ dF, double* __restrict__ out,
                           int Nx, int Ny, int Nz, double Dx, double Dy, double Dz, double Lx, double Ly, double Lz,
                           double invdV, int fixed_bnd, int central, double2* __restrict__ partials) {
  double mx[1] = {0.0};
  const long N = (long)Nx * Ny * Nz;
  const int st = central ? 2 : 1;
  for (long p = (long)blockIdx.x * kThreads + threadIdx.x; p < N; p += (long)gridDim.x * kThreads) {
    const int z = p % Nz, y = (p / Nz) % Ny, x = p / ((long)Nz * Ny);
    const long px = ((long)((x + 1) % Nx) * Ny + y) * Nz + z, mxp = ((long)((x + Nx - 1) % Nx) * Ny + y) * Nz + z;
    const long py = ((long)x * Ny + (y + 1) % Ny) * Nz + z, my = ((long)x * Ny + (y + Ny - 1) % Ny) * Nz + z;
    const long pz = ((long)x * Ny + y) * Nz + (z + 1) % Nz, mz = ((long)x * Ny + y) * Nz + (z + Nz - 1) % Nz;
    const long qx = ((long)((x + st) % Nx) * Ny + y) * Nz + z, nxq = ((long)((x + 2 * Nx - st) % Nx) * Ny + y) * Nz + z;
    const long qy = ((long)x * Ny + (y + st) % Ny) * Nz + z, nyq = ((long)x * Ny + (y + 2 * Ny - st) % Ny) * Nz + z;
    const long qz = ((long)x * Ny + y) * Nz + (z + st) % Nz, nzq = ((long)x * Ny + y) * Nz + (z + 2 * Nz - st) % Nz;
    const double x0 = dF[p], r0 = rho[p], d0 = central ? 0.0 : r0;
    double gx = Dx * ((rho[px] + d0) * (dF[qx] - x0) - (d0 + rho[mxp]) * (x0 - dF[nxq])) +
                Dy * ((rho[py] + d0) * (dF[qy] - x0) - (d0 + rho[my]) * (x0 - dF[nyq])) +
                Dz * ((rho[pz] + d0) * (dF[qz] - x0) - (d0 + rho[mz]) * (x0 - dF[nzq]));
    if (fixed_bnd && (x == 0 || y == 0 || z == 0)) gx = 0.0;
    gx *= invdV;
    mx[0] = fmax(mx[0], fabs(gx));
    gx += -Lx * (rho[px] + rho[mxp] - 2 * r0) - Ly * (rho[py] + rho[my] - 2 * r0) - Lz * (rho[pz] + rho[mz] - 2 * r0);
    out[p] = gx;
  }
  reduce_store<0, 1>(nullptr, mx, partials);
}

// w = U0 x + U1 R0 + U2 (R1 - R0) on the half spectrum (apply_diffusion_propogator, DDFT.cpp:239-266).
__global__ void k_ddft_update(const cplx<double>* __restrict__ X, const cplx<double>* __restrict__ R0,
                              const cplx<double>* __restrict__ R1, cplx<double>* __restrict__ W,
                              const double* __restrict__ lam, const double* __restrict__ ex, Geo g, double dt) {
  const long n = (long)g.Nx * g.Ny * g.Zh;
  for (long q = (long)blockIdx.x * kThreads + threadIdx.x; q < n; q += (long)gridDim.x * kThreads) {
    const int kz = q % g.Zh, ky = (q / g.Zh) % g.Ny, kx = q / ((long)g.Zh * g.Ny);
    const long s = ((long)kx * g.Ny + ky) * g.Zp + kz;
    const double Lam = lam[kx] + lam[g.Nx + ky] + lam[g.Nx + g.Ny + kz];
    const double e = ex[kx] * ex[g.Nx + ky] * ex[g.Nx + g.Ny + kz];
    const bool k0 = q == 0;
    const double U0 = e;
    const double U1 = k0 ? dt : (e - 1) / Lam;
    const double U2 = k0 ? dt / 2 : (e - 1 - dt * Lam) / (Lam * Lam * dt);
    const cplx<double> x = X[s], r0 = R0[s], r1 = R1[s];
    W[s] = {U0 * x.x + U1 * r0.x + U2 * (r1.x - r0.x), U0 * x.y + U1 * r0.y + U2 * (r1.y - r0.y)};
  }
}

// d1 <- new; max |d1 - new| and min(new) (apply_diffusion_propogator :268-292 and the
// negativity test of DDFT::step :160).  Reductions: max dev, max(-new), sum new.  A fixed
// background keeps d1 on the boundary and leaves it out of the deviation (:276-287).
__global__ void k_ddft_accept(double* __restrict__ d1, const double* __restrict__ nw, long N, int fixed_bnd, int Ny,
                              int Nz, double2* __restrict__ partials) {
  CSum s[1] = {{0, 0}};
  double mx[2] = {0.0, -1e300};
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) {
    const long yz = (long)Ny * Nz;
    const bool keep = fixed_bnd && (i < yz || (i % yz) < Nz || i % Nz == 0);
    const double w = keep ? d1[i] : nw[i];
    if (!keep) mx[0] = fmax(mx[0], fabs(d1[i] - w));
    mx[1] = fmax(mx[1], -w);
    s[0].add(w);
    d1[i] = w;
  }
  reduce_store<1, 2>(s, mx, partials);
}

// Open-boundary DDFT (fftDiffusion's unpack, mduran/src/DDFT_IF_Open.cpp:325-339, 353-375): interior sites take
// S/norm + bg from the compact back-transform, boundary sites (ix, iy or iz = 0) the background.  Reductions:
// max |d1_old - d1_new| over the interior, max(-d1_new) over all sites.
__global__ void k_open_accept(double* __restrict__ d1, const double* __restrict__ S, int Nx, int Ny, int Nz,
                              double norm, double bg, double2* __restrict__ partials) {
  double mx[2] = {0.0, -1e300};
  const long N = (long)Nx * Ny * Nz;
  for (long p = (long)blockIdx.x * kThreads + threadIdx.x; p < N; p += (long)gridDim.x * kThreads) {
    const int iz = (int)(p % Nz), iy = (int)((p / Nz) % Ny), ix = (int)(p / ((long)Nz * Ny));
    double v = bg;
    if (ix && iy && iz) {
      v = S[(((long)ix - 1) * (Ny - 1) + iy - 1) * (Nz - 1) + iz - 1] / norm + bg;
      mx[0] = fmax(mx[0], fabs(d1[p] - v));
    }
    d1[p] = v;
    mx[1] = fmax(mx[1], -v);
  }
  reduce_store<0, 2>(nullptr, mx, partials);
}

__global__ void k_copy(const double* __restrict__ a, double* __restrict__ b, long N) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) b[i] = a[i];
}

__global__ void k_dot(const double* __restrict__ a, const double* __restrict__ b, long N, double2* __restrict__ partials) {
  CSum s[1] = {{0, 0}};
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) s[0].add(a[i] * b[i]);
  reduce_store<1, 0>(s, nullptr, partials);
}

__global__ void k_axpy(double* __restrict__ y, const double* __restrict__ x, long N, double a) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) y[i] += a * x[i];
}

__global__ void k_zero(double* __restrict__ y, long N) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) y[i] = 0.0;
}

__global__ void k_sum(const double* __restrict__ a, long N, double2* __restrict__ partials) {
  CSum s[1] = {{0, 0}};
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads) s[0].add(a[i]);
  reduce_store<1, 0>(s, nullptr, partials);
}

// Launch a grid-stride vector kernel plus its finishing reduction; results to host.
struct Reducer {
  DevBuf part, out;
  void init() {
    part.alloc(sizeof(double2) * kRedBlocks * 8);
    out.alloc(sizeof(double) * 8);
  }
  void finish(hipStream_t st, int NS, int NM, double* host) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kThreads), 0, st, part.as<double2>(), kRedBlocks, NS, NM, out.as<double>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, out.p, sizeof(double) * (NS + NM), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
};

// Every reducing kernel runs exactly kRedBlocks grid-stride blocks so that all
// kRedBlocks partial slots are written (idle threads contribute the identity).
dim3 vgrid(long) { return dim3((unsigned)kRedBlocks); }

}  // namespace
}  // namespace cdft_impl

using namespace cdft_impl;

// =============================== FIRE2 =====================================
struct cdft_fire_s {
  cdft_system sys = nullptr;    // the evaluating system (a mixture's first member, for its stream)
  cdft_mixture mix = nullptr;   // several species: x, v, dF, rho hold them one after the other
  long N = 0;                   // elements of the concatenated variables (species x lattice sites)
  double dV = 0;
  AliasSet al{};
  DevBuf x, v, dF, raw, rho, x_rem, v_rem, dF_rem, raw_rem;
  DevBuf dir;  // Minimizer::fixed_direction_ (allocated when set)
  bool has_dir = false;
  Reducer red;
  cdft_fire_params p{};
  // reference state (Minimizer.h:45-56, 105-160)
  double dt = 1e-3, dt_max = 1e-2, dt_best = 0, alpha = 0.1;
  int N_P_positive = 0, N_P_negative = 0, backtracks = 0;
  unsigned it = 0;
  double vv = 1.0, F = 0, fold = 0, fmax = -1, fmax_raw = 0, rms_force = 0, vnorm = 0, P_next = 0;
  int calls = 0, steps = 0;
  double parts[5] = {0, 0, 0, 0, 0};

  void launch_post() {
    const dim3 g = vgrid(N);
    hipLaunchKernelGGL(k_fire_post, g, dim3(kThreads), 0, sys->stream, x.as<double>(), v.as<double>(),
                       raw.as<double>(), dF.as<double>(), N, al, red.part.as<double2>());
    HIPCHK(hipGetLastError());
  }
  // Minimizer::getDF_DX (Minimizer.cpp:90-115) from the current rho buffer.
  int getDF_DX() {
    int rc;
    if (mix) {
      const int S = mixture_species(*mix);
      std::vector<const double*> r(S);
      std::vector<double*> d(S);
      for (int i = 0; i < S; ++i) {
        r[i] = rho.as<double>() + i * al.nseg;
        d[i] = raw.as<double>() + i * al.nseg;
      }
      rc = mixture_eval(*mix, r.data(), d.data(), false, parts, true, false);
    } else {
      enqueue_eval(sys, rho.as<double>(), raw.as<double>(), true, false);
      rc = fetch_parts(sys, parts);
    }
    if (rc == CDFT_ETA_TOO_LARGE) return rc;
    calls++;
    F = parts[0];
    if (has_dir) {  // Minimizer.cpp:103-110: dF -= 2 (e . dF) e
      hipLaunchKernelGGL(k_dot, vgrid(N), dim3(kThreads), 0, sys->stream, dir.as<double>(), raw.as<double>(), N,
                         red.part.as<double2>());
      HIPCHK(hipGetLastError());
      double d;
      red.finish(sys->stream, 1, 0, &d);
      hipLaunchKernelGGL(k_axpy, vgrid(N), dim3(kThreads), 0, sys->stream, raw.as<double>(), dir.as<double>(), N,
                         -2 * d);
      HIPCHK(hipGetLastError());
    }
    launch_post();
    double r[4];
    red.finish(sys->stream, 2, 2, r);
    rms_force = std::sqrt(r[0] / N);
    P_next = -r[1];
    fmax = r[2] / dV;
    fmax_raw = r[3] / dV;
    return CDFT_OK;
  }
  // fireMinimizer2::SemiImplicitEuler (Minimizer.cpp:273-327)
  int semi_implicit_euler() {
    const dim3 g = vgrid(N);
    hipLaunchKernelGGL(k_fire_kick, g, dim3(kThreads), 0, sys->stream, v.as<double>(), dF.as<double>(), N, dt,
                       red.part.as<double2>());
    HIPCHK(hipGetLastError());
    double r[2];
    red.finish(sys->stream, 2, 0, r);
    const double vn2 = r[0], fn2 = r[1];
    rms_force = std::sqrt(fn2 / N);
    vnorm = std::sqrt(vn2 / N);
    const double mix = fn2 > 0 ? -alpha * std::sqrt(vn2 / fn2) : 0.0;
    hipLaunchKernelGGL(k_fire_mix, g, dim3(kThreads), 0, sys->stream, x.as<double>(), v.as<double>(), dF.as<double>(),
                       rho.as<double>(), N, alpha, mix, dt, al);
    HIPCHK(hipGetLastError());
    return getDF_DX();
  }
  // fireMinimizer2::step (Minimizer.cpp:160-271); snapshot I/O omitted.
  int step() {
    it++;
    steps++;
    const double P = P_next;
    fold = F;
    const dim3 g = vgrid(N);
    hipLaunchKernelGGL(k_copy4, g, dim3(kThreads), 0, sys->stream, x.as<double>(), v.as<double>(), dF.as<double>(),
                       raw.as<double>(), x_rem.as<double>(), v_rem.as<double>(), dF_rem.as<double>(),
                       raw_rem.as<double>(), N, 1.0);
    HIPCHK(hipGetLastError());
    if (P > 0) {
      N_P_positive++;
      N_P_negative = 0;
      if (N_P_positive > p.N_delay) {
        dt = std::min(dt * p.f_inc, dt_max);
        alpha = alpha * p.f_alf;
      }
    } else {
      N_P_positive = 0;
      N_P_negative++;
      if (N_P_negative > p.N_P_negative_max) {
        g_err = "Cannot stop going uphill in Fire2";
        return CDFT_UPHILL_LIMIT;
      }
      if (!(p.initial_delay && it < (unsigned)p.N_delay)) {
        if (dt * p.f_dec >= p.dt_min) dt *= p.f_dec;
        alpha = p.alpha_start;
      }
      hipLaunchKernelGGL(k_fire_uphill, g, dim3(kThreads), 0, sys->stream, x.as<double>(), v.as<double>(), N, dt);
      HIPCHK(hipGetLastError());
      backtracks++;
    }
    int status = semi_implicit_euler();
    if (status == CDFT_OK) {
      vv = vv * 0.9 + 0.1 * (std::fabs(F - fold) / dt);
      dt_best = std::max(dt, dt_best);
    } else {
      // restore x, v/2, dF (Minimizer.cpp:233-245)
      hipLaunchKernelGGL(k_copy4, g, dim3(kThreads), 0, sys->stream, x_rem.as<double>(), v_rem.as<double>(),
                         dF_rem.as<double>(), raw_rem.as<double>(), x.as<double>(), v.as<double>(), dF.as<double>(),
                         raw.as<double>(), N, 0.5);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(sys->stream));
      dt /= 2;
      fmax = 1000;
      backtracks++;
      P_next = 0.5 * P;  // -(v/2).dF of the restored state
      F = fold;
    }
    if (backtracks >= 10) {
      dt_max = std::min(dt_best, 0.9 * dt_max);
      dt_best = 0;
      backtracks = 0;
    }
    return status == CDFT_OK ? CDFT_OK : CDFT_ETA_TOO_LARGE;
  }
  void report(cdft_fire_report* r) const {
    if (!r) return;
    r->F = F;
    r->fmax_alias = fmax;
    r->fmax_raw = fmax_raw;
    r->dt = dt;
    r->alpha = alpha;
    r->vv = vv;
    r->rms_force = rms_force;
    r->steps = steps;
    r->calls = calls;
    r->monitor = std::fabs(vv / (al.nseg * dV));  // |vv| / the lattice volume (Minimizer.h:105), any species count
  }
};

// =============================== DDFT ======================================
struct cdft_ddft_s {
  cdft_system sys = nullptr;
  long N = 0;
  bool central = false;  // central differences in g . dF (DDFT(dft, showGraphics, central_differences))
  double dx = 0, dy = 0, dz = 0;
  double dt = 0, time = 0, tol = 1e-4, dtMax = 1.0, RHS_max = 1.0, F = 0, mass = 0;
  int successes = 0, calls = 0, restarts = 0;
  std::vector<double> Lx, Ly, Lz;
  DevBuf d0, d1, dF, tmp, X, R0, R1, W, lam, ex;
  DevBuf sp;  // the species density of an evaluation inside the fixed-point loop (fixed mass rescales it)
  Reducer red;
  double parts[5] = {0, 0, 0, 0, 0};
  bool ex_ready = false;

  void r2c(const double* in, cplx<double>* out) {
    launch_rows_r2c<double, double>(*sys, in, 0, out, 0, 1);
    launch_cols<double>(*sys, out, 0, 1, 1, true);
    launch_cols<double>(*sys, out, 0, 1, 0, true);
  }
  void c2r(cplx<double>* in, double* out) {  // destroys `in`
    launch_cols<double>(*sys, in, 0, 1, 0, false);
    launch_cols<double>(*sys, in, 0, 1, 1, false);
    launch_rows_c2r<double, double>(*sys, in, 0, out, 0, 1, 1.0 / (double)N);
  }
  int eval(const double* rho) {
    calls++;
    enqueue_eval(sys, rho, dF.as<double>(), true, false);
    const int rc = fetch_parts(sys, parts);
    if (rc == CDFT_OK) F = parts[0];
    return rc;
  }
  // DDFT::change_timestep (DDFT.cpp:184-207)
  void change_timestep(double ndt) {
    if (std::fabs(dt - ndt) < 1e-10 && ex_ready) return;
    dt = ndt;
    std::vector<double> e;
    for (double l : Lx) e.push_back(std::exp(l * dt));
    for (double l : Ly) e.push_back(std::exp(l * dt));
    for (double l : Lz) e.push_back(std::exp(l * dt));
    HIPCHK(hipMemcpyAsync(ex.p, e.data(), sizeof(double) * e.size(), hipMemcpyHostToDevice, sys->stream));
    HIPCHK(hipStreamSynchronize(sys->stream));
    ex_ready = true;
  }
  double excess_rhs(const double* rho, cplx<double>* R) {
    const double st = central ? 2.0 : 1.0;  // DDFT::g_dot_x stride, src/DDFT.cpp:325-327
    const double D[3] = {1 / (2 * st * dx * dx), 1 / (2 * st * dy * dy), 1 / (2 * st * dz * dz)};
    const double L[3] = {1 / (dx * dx), 1 / (dy * dy), 1 / (dz * dz)};
    const Geo& g = sys->g;
    hipLaunchKernelGGL(k_ddft_rhs, vgrid(N), dim3(kThreads), 0, sys->stream, rho, dF.as<double>(), tmp.as<double>(),
                       g.Nx, g.Ny, g.Nz, D[0], D[1], D[2], L[0], L[1], L[2], 1.0 / (dx * dy * dz),
                       (sys->d.boundary & CDFT_FIXED_BACKGROUND) ? 1 : 0, central ? 1 : 0, red.part.as<double2>());
    HIPCHK(hipGetLastError());
    r2c(tmp.as<double>(), R);
    double r;
    red.finish(sys->stream, 0, 1, &r);
    return r / (dx * dy * dz);
  }
  // apply_diffusion_propogator (DDFT.cpp:239-292): X holds the spectrum of d0.
  void propagate(double& maxdev, double& minval, double& sum) {
    const Geo& g = sys->g;
    hipLaunchKernelGGL(k_ddft_update, vgrid((long)g.Nx * g.Ny * g.Zh), dim3(kThreads), 0, sys->stream,
                       X.as<cplx<double>>(), R0.as<cplx<double>>(), R1.as<cplx<double>>(), W.as<cplx<double>>(),
                       lam.as<double>(), ex.as<double>(), g, dt);
    HIPCHK(hipGetLastError());
    c2r(W.as<cplx<double>>(), tmp.as<double>());
    hipLaunchKernelGGL(k_ddft_accept, vgrid(N), dim3(kThreads), 0, sys->stream, d1.as<double>(), tmp.as<double>(), N,
                       (sys->d.boundary & CDFT_FIXED_BACKGROUND) ? 1 : 0, g.Ny, g.Nz, red.part.as<double2>());
    HIPCHK(hipGetLastError());
    double r[3];
    red.finish(sys->stream, 1, 2, r);
    sum = r[0];
    maxdev = r[1];
    minval = -r[2];
  }
  void copy(const DevBuf& a, DevBuf& b) {
    hipLaunchKernelGGL(k_copy, vgrid(N), dim3(kThreads), 0, sys->stream, a.as<double>(), b.as<double>(), N);
    HIPCHK(hipGetLastError());
  }
  // DDFT::step (DDFT.cpp:113-182)
  int step() {
    if (!ex_ready) { const double t = dt; dt = 0; change_timestep(t); }
    int rc = eval(d0.as<double>());
    if (rc) return rc;
    copy(d0, d1);
    RHS_max = excess_rhs(d0.as<double>(), R0.as<cplx<double>>());
    r2c(d0.as<double>(), X.as<cplx<double>>());
    bool restart, decreased = false;
    double sum = 0.0;
    do {
      double dev = 1;
      restart = false;
      for (int i = 0; i < 100 && dev > tol && !restart; ++i) {
        // species->set_density(d1): with a fixed mass the evaluation rescales the species copy, not d1
        const bool copy_sp = sys->d.fixed_mass > 0;
        if (copy_sp) copy(d1, sp);
        const double* cur = copy_sp ? sp.as<double>() : d1.as<double>();
        rc = eval(cur);
        if (rc) return rc;
        excess_rhs(cur, R1.as<cplx<double>>());
        const double old = dev;
        double mn;
        propagate(dev, mn, sum);
        if (mn < 0 || (i > 0 && old < dev)) restart = true;
      }
      if (restart || dev > tol) {
        const double before = dt;
        change_timestep(dt / 10);
        if (dt == before) {  // DDFT.cpp:184-186 ignores the change: the same attempt would repeat forever
          g_err = "DDFT: the fixed-point iteration fails at the smallest time step change_timestep can set";
          return CDFT_DDFT_STALLED;
        }
        copy(d0, d1);
        decreased = true;
        restart = true;
        restarts++;
      }
    } while (restart);
    time += dt;
    if (decreased) successes = 0;
    else successes++;
    if (successes >= 5 && dt < dtMax) {
      change_timestep(std::min(2 * dt, dtMax));
      successes = 0;
    }
    copy(d1, d0);
    mass = sum * dx * dy * dz;
    return eval(d0.as<double>());
  }
};

// ============================ open-boundary DDFT ============================
// DDFT_IF_Open (mduran/src/DDFT_IF_Open.cpp:25-463): the density is held at `bg` on the planes ix, iy or
// iz = 0 and the interior evolves with the integrating factor of the Dirichlet Laplacian (3D DST-I,
// dst.hpp).  step() = step_string (:94-167) + the time-step bookkeeping of step (:171-184), with the
// reference's control flow: at most 20 fixed-point iterations, dt halved on restart and restored after.
struct cdft_ddft_open_s {
  cdft_system sys = nullptr;
  long N = 0, ni = 0;
  int n[3] = {0, 0, 0};  // interior extents Nx-1, Ny-1, Nz-1
  double dx = 0, dy = 0, dz = 0, bg = 0;
  double dt = 0, dtMax = 1.0, tol = 1e-4, F = 0, time = 0, rhs_max = 0;
  int successes = 0, calls = 0, restarts = 0, evals = 0;
  std::vector<double> lam;  // Lx | Ly | Lz (:58-82)
  DevBuf d0, d1, dF, tmp, S, R0s, R1s, lamd, ex;
  Radices rad[3];
  DevBuf tw[3];
  int nbl[3] = {1, 1, 1};
  Reducer red;
  double parts[5] = {0, 0, 0, 0, 0};

  int eval(const double* rho) {
    evals++;
    enqueue_eval(sys, rho, dF.as<double>(), true, false);
    const int rc = fetch_parts(sys, parts);
    if (rc == CDFT_OK) F = parts[0];
    return rc;
  }
  void dst3(double* a) {
    for (int ax = 2; ax >= 0; --ax) {
      const long nlines = ni / n[ax];
      hipLaunchKernelGGL(k_dst1, dim3(div_up(nlines, nbl[ax])), dim3(256),
                         sizeof(cplx<double>) * (size_t)nbl[ax] * rad[ax].L, sys->stream, a, n[0], n[1], n[2], ax,
                         rad[ax], tw[ax].as<cplx<double>>(), nbl[ax]);
      HIPCHK(hipGetLastError());
    }
  }
  void pack(const double* x, double val, double* o) {
    hipLaunchKernelGGL(k_open_pack, vgrid(ni), dim3(kThreads), 0, sys->stream, x, o, sys->g.Nx, sys->g.Ny,
                       sys->g.Nz, val);
    HIPCHK(hipGetLastError());
  }
  // calcNonlinearTerm (:377-455) of d with the current dF, packed and transformed
  void rhs_sin(const double* d, DevBuf& out) {
    const Geo& g = sys->g;
    hipLaunchKernelGGL(k_ddft_rhs, vgrid(N), dim3(kThreads), 0, sys->stream, d, dF.as<double>(), tmp.as<double>(),
                       g.Nx, g.Ny, g.Nz, 1 / (2 * dx * dx), 1 / (2 * dy * dy), 1 / (2 * dz * dz), 1 / (dx * dx),
                       1 / (dy * dy), 1 / (dz * dz), 1.0 / (dx * dy * dz), 0, 0, red.part.as<double2>());
    HIPCHK(hipGetLastError());
    pack(tmp.as<double>(), 0.0, out.as<double>());
    dst3(out.as<double>());
  }
  void copy(const DevBuf& a, DevBuf& b, long n_) {
    hipLaunchKernelGGL(k_copy, vgrid(n_), dim3(kThreads), 0, sys->stream, a.as<double>(), b.as<double>(), n_);
    HIPCHK(hipGetLastError());
  }
  // fftDiffusion (:268-339) from d0 into d1; returns the max deviation, sets the minimum of the new d1
  double diffuse(double dtl, double& mn) {
    pack(d0.as<double>(), bg, S.as<double>());
    dst3(S.as<double>());
    std::vector<double> e(lam.size());
    for (size_t i = 0; i < lam.size(); ++i) e[i] = std::exp(lam[i] * dtl);
    HIPCHK(hipMemcpyAsync(ex.p, e.data(), sizeof(double) * e.size(), hipMemcpyHostToDevice, sys->stream));
    hipLaunchKernelGGL(k_open_update, vgrid(ni), dim3(kThreads), 0, sys->stream, S.as<double>(), R0s.as<double>(),
                       R1s.as<double>(), lamd.as<double>(), ex.as<double>(), n[0], n[1], n[2], dtl);
    HIPCHK(hipGetLastError());
    dst3(S.as<double>());
    const Geo& g = sys->g;
    hipLaunchKernelGGL(k_open_accept, vgrid(N), dim3(kThreads), 0, sys->stream, d1.as<double>(), S.as<double>(),
                       g.Nx, g.Ny, g.Nz, 8.0 * g.Nx * g.Ny * g.Nz, bg, red.part.as<double2>());
    HIPCHK(hipGetLastError());
    double r[2];
    red.finish(sys->stream, 0, 2, r);  // synchronises: e stays alive until the copy is done
    mn = -r[1];
    return r[0];
  }
  int step() {
    double dtl = dt;
    int rc = eval(d0.as<double>());
    if (rc) return rc;
    rhs_sin(d0.as<double>(), R0s);
    {
      double r;
      red.finish(sys->stream, 0, 1, &r);  // max |g . dF| / dV of k_ddft_rhs
      rhs_max = r / (dx * dy * dz);
    }
    copy(d0, d1, N);
    copy(R0s, R1s, ni);
    double deviation = 1;
    bool restart;
    int halvings = 0;
    do {
      restart = false;
      double old_error = 0;
      for (int i = 0; i < 20 && deviation > tol && !restart; ++i) {
        double mn;
        deviation = diffuse(dtl, mn);
        if (mn < 0 || (i > 0 && old_error < deviation)) restart = true;
        else {
          rc = eval(d1.as<double>());
          if (rc) return rc;
          rhs_sin(d1.as<double>(), R1s);
        }
        old_error = deviation;
      }
      if (deviation > tol) restart = true;
      if (restart) {
        REQUIRE(++halvings <= 60, "open DDFT: the time step could not be reduced enough");
        dtl /= 2;
        restarts++;
        copy(d0, d1, N);
        copy(R0s, R1s, ni);
      }
    } while (restart);
    copy(d1, d0, N);
    HIPCHK(hipStreamSynchronize(sys->stream));
    calls++;
    time += dtl;
    if (dtl < dt * 1.01) successes = 0;
    else successes++;
    if (successes >= 5 && dt < dtMax) {
      dt = std::min(2 * dtl, dtMax);
      successes = 0;
    }
    return CDFT_OK;
  }
};

namespace {
void default_fire_params(cdft_fire_params* p) {  // fireMinimizer2 constructor, Minimizer.cpp:117-147
  p->dt = 1e-3;
  p->dt_max = 1e-2;
  p->dt_min = 0.0;
  p->alpha_start = 0.1;
  p->f_dec = 0.5;
  p->f_inc = 1.1;
  p->f_alf = 0.99;
  p->N_delay = 5;
  p->N_P_negative_max = 20;
  p->initial_delay = 1;
}
}  // namespace

extern "C" {

int cdft_fire_default_params(cdft_fire_params* p) {
  return guarded([&] {
    REQUIRE(p, "null argument");
    default_fire_params(p);
    return CDFT_OK;
  });
}

int cdft_fire_create(cdft_system sys, const double* rho0, int flags, const cdft_fire_params* params, cdft_fire* out) {
  return guarded([&] {
    REQUIRE(sys && rho0 && out, "null argument");
    REQUIRE(sys->nb == 1, "use cdft_fire_batch_create for batched systems");
    auto f = std::make_unique<cdft_fire_s>();
    f->sys = sys;
    f->N = sys->N;
    f->dV = sys->d.dx * sys->d.dy * sys->d.dz;
    if (params) f->p = *params;
    else default_fire_params(&f->p);
    f->dt = f->p.dt;
    f->dt_max = f->p.dt_max;
    f->dt_best = f->dt;
    f->alpha = f->p.alpha_start;
    f->al.nseg = f->N;
    f->al.ns = 1;
    f->al.a[0].fmt = sys->m > 0;
    const double d = sys->d.hsd;
    f->al.a[0].c = (1.0 - 1e-18 * (4 * M_PI * d * d * d / 3)) / f->dV;
    const size_t b = sizeof(double) * f->N;
    for (DevBuf* q : {&f->x, &f->v, &f->dF, &f->raw, &f->rho, &f->x_rem, &f->v_rem, &f->dF_rem, &f->raw_rem}) q->alloc(b);
    f->red.init();
    HIPCHK(hipMemcpyAsync(f->rho.p, rho0, b, (flags & CDFT_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                          sys->stream));
    HIPCHK(hipMemsetAsync(f->v.p, 0, b, sys->stream));
    hipLaunchKernelGGL(k_alias_from_rho, vgrid(f->N), dim3(kThreads), 0, sys->stream, f->rho.as<double>(),
                       f->x.as<double>(), f->N, f->al);
    hipLaunchKernelGGL(k_rho_from_alias, vgrid(f->N), dim3(kThreads), 0, sys->stream, f->x.as<double>(),
                       f->rho.as<double>(), f->N, f->al);
    HIPCHK(hipGetLastError());
    const int rc = f->getDF_DX();
    if (rc != CDFT_OK) return rc;
    *out = f.release();
    return (int)CDFT_OK;
  });
}

int cdft_mixture_fire_create(cdft_mixture mix, const double* const* rho0, int flags, const cdft_fire_params* params,
                             cdft_fire* out) {
  return guarded([&] {
    REQUIRE(mix && rho0 && out, "null argument");
    const int S = mixture_species(*mix);
    REQUIRE(S <= kMaxAliasSeg, "the mixture's FIRE takes at most 8 species");
    for (int i = 0; i < S; ++i) REQUIRE(rho0[i], "null species density");
    cdft_system s0 = mixture_member(*mix, 0);
    auto f = std::make_unique<cdft_fire_s>();
    f->sys = s0;
    f->mix = mix;
    f->N = s0->N * S;
    f->dV = s0->d.dx * s0->d.dy * s0->d.dz;
    if (params) f->p = *params;
    else default_fire_params(&f->p);
    f->dt = f->p.dt;
    f->dt_max = f->p.dt_max;
    f->dt_best = f->dt;
    f->alpha = f->p.alpha_start;
    f->al.nseg = s0->N;
    f->al.ns = S;
    for (int i = 0; i < S; ++i) {  // FMT_Species aliases, each with its own diameter
      const double d = mixture_member(*mix, i)->d.hsd;
      f->al.a[i].fmt = 1;
      f->al.a[i].c = (1.0 - 1e-18 * (4 * M_PI * d * d * d / 3)) / f->dV;
    }
    const size_t b = sizeof(double) * f->N;
    for (DevBuf* q : {&f->x, &f->v, &f->dF, &f->raw, &f->rho, &f->x_rem, &f->v_rem, &f->dF_rem, &f->raw_rem}) q->alloc(b);
    f->red.init();
    for (int i = 0; i < S; ++i)
      HIPCHK(hipMemcpyAsync(f->rho.as<double>() + i * s0->N, rho0[i], sizeof(double) * s0->N,
                            (flags & CDFT_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s0->stream));
    HIPCHK(hipMemsetAsync(f->v.p, 0, b, s0->stream));
    hipLaunchKernelGGL(k_alias_from_rho, vgrid(f->N), dim3(kThreads), 0, s0->stream, f->rho.as<double>(),
                       f->x.as<double>(), f->N, f->al);
    hipLaunchKernelGGL(k_rho_from_alias, vgrid(f->N), dim3(kThreads), 0, s0->stream, f->x.as<double>(),
                       f->rho.as<double>(), f->N, f->al);
    HIPCHK(hipGetLastError());
    const int rc = f->getDF_DX();
    if (rc != CDFT_OK) return rc;
    *out = f.release();
    return (int)CDFT_OK;
  });
}

int cdft_fire_set_fixed_direction(cdft_fire f, const double* dir, int flags) {
  return guarded([&] {
    REQUIRE(f, "null argument");
    if (!dir) {
      f->has_dir = false;
      return (int)CDFT_OK;
    }
    const size_t b = sizeof(double) * f->N;
    if (!f->dir.p) f->dir.alloc(b);
    HIPCHK(hipMemcpyAsync(f->dir.p, dir, b, (flags & CDFT_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                          f->sys->stream));
    HIPCHK(hipStreamSynchronize(f->sys->stream));
    f->has_dir = true;
    return (int)CDFT_OK;
  });
}

// fireMinimizer2::reset (Minimizer.cpp:30-34, 149-158): aliases from rho, dt_best = dt, getDF_DX, v = 0; dt, alpha
// and the step counters are kept.
int cdft_fire_reset(cdft_fire f, const double* rho, int flags) {
  return guarded([&] {
    REQUIRE(f && rho, "null argument");
    hipStream_t st = f->sys->stream;
    HIPCHK(hipMemcpyAsync(f->rho.p, rho, sizeof(double) * f->N,
                          (flags & CDFT_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_alias_from_rho, vgrid(f->N), dim3(kThreads), 0, st, f->rho.as<double>(), f->x.as<double>(),
                       f->N, f->al);
    hipLaunchKernelGGL(k_rho_from_alias, vgrid(f->N), dim3(kThreads), 0, st, f->x.as<double>(), f->rho.as<double>(),
                       f->N, f->al);
    HIPCHK(hipGetLastError());
    f->dt_best = f->dt;
    const int rc = f->getDF_DX();
    hipLaunchKernelGGL(k_zero, vgrid(f->N), dim3(kThreads), 0, st, f->v.as<double>(), f->N);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    f->P_next = 0.0;  // P = -v . dF of the next step, with v = 0
    return rc;
  });
}

int cdft_fire_step(cdft_fire f, cdft_fire_report* rep) {
  return guarded([&] {
    REQUIRE(f, "null argument");
    const int rc = f->step();
    f->report(rep);
    return rc;
  });
}

int cdft_fire_run(cdft_fire f, int max_steps, double tol_raw, cdft_fire_report* rep) {
  return guarded([&] {
    REQUIRE(f, "null argument");
    int rc = CDFT_OK;
    for (int s = 0; s < max_steps; ++s) {
      rc = f->step();
      if (rc == CDFT_UPHILL_LIMIT) break;
      rc = CDFT_OK;  // a back-tracked step is part of the normal iteration
      if (f->fmax_raw <= tol_raw && f->fmax >= 0 && f->fmax < 1000) break;
    }
    f->report(rep);
    return rc;
  });
}

int cdft_fire_get(cdft_fire f, double* rho, double* dF_raw, double* x, double* v) {
  return guarded([&] {
    REQUIRE(f, "null argument");
    const size_t b = sizeof(double) * f->N;
    hipLaunchKernelGGL(k_rho_from_alias, vgrid(f->N), dim3(kThreads), 0, f->sys->stream, f->x.as<double>(),
                       f->rho.as<double>(), f->N, f->al);
    HIPCHK(hipGetLastError());
    // with a fixed mass: the density the next evaluation uses
    if (!f->mix) constraints_pre(*f->sys, f->rho.as<double>());
    else
      for (int i = 0; i < mixture_species(*f->mix); ++i)
        constraints_pre(*mixture_member(*f->mix, i), f->rho.as<double>() + i * f->al.nseg);
    if (rho) HIPCHK(hipMemcpyAsync(rho, f->rho.p, b, hipMemcpyDeviceToHost, f->sys->stream));
    if (dF_raw) HIPCHK(hipMemcpyAsync(dF_raw, f->raw.p, b, hipMemcpyDeviceToHost, f->sys->stream));
    if (x) HIPCHK(hipMemcpyAsync(x, f->x.p, b, hipMemcpyDeviceToHost, f->sys->stream));
    if (v) HIPCHK(hipMemcpyAsync(v, f->v.p, b, hipMemcpyDeviceToHost, f->sys->stream));
    HIPCHK(hipStreamSynchronize(f->sys->stream));
    return CDFT_OK;
  });
}

int cdft_fire_destroy(cdft_fire f) {
  return guarded([&] {
    delete f;
    return CDFT_OK;
  });
}

int cdft_ddft_create(cdft_system sys, const double* rho0, double dt, double dt_max, double tol, cdft_ddft* out) {
  return guarded([&] {
    REQUIRE(sys && rho0 && out, "null argument");
    REQUIRE(sys->nb == 1, "DDFT runs on single systems");
    auto d = std::make_unique<cdft_ddft_s>();
    d->sys = sys;
    d->N = sys->N;
    d->dx = sys->d.dx;
    d->dy = sys->d.dy;
    d->dz = sys->d.dz;
    d->dt = dt > 0 ? dt : 0.0001 * d->dx * d->dx;  // DDFT.cpp:46
    d->dtMax = dt_max > 0 ? dt_max : 1.0;
    d->tol = tol != 0 ? tol : 1e-4;  // any other value as DDFT::set_tolerence_fixed_point takes it
    const size_t b = sizeof(double) * d->N;
    const size_t cb = sizeof(cplx<double>) * sys->Nkp;
    for (DevBuf* q : {&d->d0, &d->d1, &d->dF, &d->tmp}) q->alloc(b);
    if (sys->d.fixed_mass > 0) d->sp.alloc(b);
    for (DevBuf* q : {&d->X, &d->R0, &d->R1, &d->W}) {
      q->alloc(cb);
      HIPCHK(hipMemset(q->p, 0, cb));
    }
    d->red.init();
    const Geo& g = sys->g;
    const double Dx = 1 / (d->dx * d->dx), Dy = 1 / (d->dy * d->dy), Dz = 1 / (d->dz * d->dz);
    for (int i = 0; i < g.Nx; ++i) d->Lx.push_back(2 * Dx * (std::cos((2 * M_PI * i) / g.Nx) - 1));  // :70-92
    for (int i = 0; i < g.Ny; ++i) d->Ly.push_back(2 * Dy * (std::cos((2 * M_PI * i) / g.Ny) - 1));
    for (int i = 0; i < g.Zh; ++i) d->Lz.push_back(2 * Dz * (std::cos((2 * M_PI * i) / g.Nz) - 1));
    std::vector<double> lam(d->Lx);
    lam.insert(lam.end(), d->Ly.begin(), d->Ly.end());
    lam.insert(lam.end(), d->Lz.begin(), d->Lz.end());
    d->lam.alloc(sizeof(double) * lam.size());
    d->ex.alloc(sizeof(double) * lam.size());
    HIPCHK(hipMemcpy(d->lam.p, lam.data(), sizeof(double) * lam.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d->d0.p, rho0, b, hipMemcpyHostToDevice));
    const int rc = d->eval(d->d0.as<double>());  // DDFT.cpp:49
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum, vgrid(d->N), dim3(kThreads), 0, sys->stream, d->d0.as<double>(), d->N,
                       d->red.part.as<double2>());
    double s;
    d->red.finish(sys->stream, 1, 0, &s);
    d->mass = s * d->dx * d->dy * d->dz;
    *out = d.release();
    return (int)CDFT_OK;
  });
}

int cdft_ddft_step(cdft_ddft d, cdft_ddft_report* rep) {
  return guarded([&] {
    REQUIRE(d, "null argument");
    const int rc = d->step();
    if (rep) {
      rep->F = d->F;
      rep->dt = d->dt;
      rep->time = d->time;
      rep->rhs_max = d->RHS_max;
      rep->calls = d->calls;
      rep->restarts = d->restarts;
      rep->mass = d->mass;
    }
    return rc;
  });
}

int cdft_ddft_set_max_time_step(cdft_ddft d, double dt_max) {
  return guarded([&] {
    REQUIRE(d && dt_max > 0, "bad argument");
    d->dtMax = dt_max;
    return CDFT_OK;
  });
}

int cdft_ddft_set_central_differences(cdft_ddft d, int on) {
  return guarded([&] {
    REQUIRE(d, "null argument");
    d->central = on != 0;
    return CDFT_OK;
  });
}

int cdft_ddft_get_density(cdft_ddft d, double* rho) {
  return guarded([&] {
    REQUIRE(d && rho, "null argument");
    HIPCHK(hipMemcpyAsync(rho, d->d0.p, sizeof(double) * d->N, hipMemcpyDeviceToHost, d->sys->stream));
    HIPCHK(hipStreamSynchronize(d->sys->stream));
    return CDFT_OK;
  });
}

int cdft_ddft_destroy(cdft_ddft d) {
  return guarded([&] {
    delete d;
    return CDFT_OK;
  });
}

int cdft_ddft_open_create(cdft_system sys, const double* rho0, double background, double dt, double dt_max, double tol,
                          cdft_ddft_open* out) {
  return guarded([&] {
    REQUIRE(sys && rho0 && out, "null argument");
    REQUIRE(sys->nb == 1, "the open DDFT runs on single systems");
    REQUIRE(!(sys->d.fixed_mass > 0) && sys->d.boundary == 0,
            "the open DDFT fixes the boundary itself: no species constraints");
    const Geo& g = sys->g;
    REQUIRE(g.Nx > 1 && g.Ny > 1 && g.Nz > 1, "the open DDFT needs at least two points per axis");
    for (int L : {2 * g.Nx, 2 * g.Ny, 2 * g.Nz})
      REQUIRE(smooth23(L) && L <= 1024, "the open DDFT needs Nx, Ny, Nz whose doubles are products of 2 and 3, <= 1024");
    auto d = std::make_unique<cdft_ddft_open_s>();
    d->sys = sys;
    d->N = sys->N;
    d->n[0] = g.Nx - 1;
    d->n[1] = g.Ny - 1;
    d->n[2] = g.Nz - 1;
    d->ni = (long)d->n[0] * d->n[1] * d->n[2];
    d->dx = sys->d.dx;
    d->dy = sys->d.dy;
    d->dz = sys->d.dz;
    d->bg = background;
    d->dt = dt > 0 ? dt : 0.0001 * d->dx * d->dx;
    d->dtMax = dt_max > 0 ? dt_max : 1.0;
    d->tol = tol != 0 ? tol : 1e-4;  // any other value as DDFT::set_tolerence_fixed_point takes it
    const int Ns[3] = {g.Nx, g.Ny, g.Nz};
    const double Ds[3] = {1 / (d->dx * d->dx), 1 / (d->dy * d->dy), 1 / (d->dz * d->dz)};
    for (int ax = 0; ax < 3; ++ax) {
      for (int i = 0; i < Ns[ax] - 1; ++i) d->lam.push_back(2 * Ds[ax] * (std::cos((M_PI * (i + 1)) / Ns[ax]) - 1));
      d->rad[ax] = make_radices(2 * Ns[ax]);
      upload_twiddles<double>(d->tw[ax], 2 * Ns[ax]);
      d->nbl[ax] = std::max(1, (256 * kDstPts) / (2 * Ns[ax]));
    }
    const size_t b = sizeof(double) * d->N, bi = sizeof(double) * d->ni;
    for (DevBuf* q : {&d->d0, &d->d1, &d->dF, &d->tmp}) q->alloc(b);
    for (DevBuf* q : {&d->S, &d->R0s, &d->R1s}) q->alloc(bi);
    d->lamd.alloc(sizeof(double) * d->lam.size());
    d->ex.alloc(sizeof(double) * d->lam.size());
    HIPCHK(hipMemcpy(d->lamd.p, d->lam.data(), sizeof(double) * d->lam.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d->d0.p, rho0, b, hipMemcpyHostToDevice));
    d->red.init();
    const int rc = d->eval(d->d0.as<double>());
    if (rc) return rc;
    *out = d.release();
    return (int)CDFT_OK;
  });
}

int cdft_ddft_open_step(cdft_ddft_open d, cdft_ddft_open_report* rep) {
  return guarded([&] {
    REQUIRE(d, "null argument");
    const int rc = d->step();
    if (rep) {
      rep->F = d->F;
      rep->dt = d->dt;
      rep->time = d->time;
      rep->rhs_max = d->rhs_max;
      rep->calls = d->calls;
      rep->restarts = d->restarts;
      rep->evaluations = d->evals;
    }
    return rc;
  });
}

int cdft_ddft_open_get_density(cdft_ddft_open d, double* rho) {
  return guarded([&] {
    REQUIRE(d && rho, "null argument");
    HIPCHK(hipMemcpy(rho, d->d0.p, sizeof(double) * d->N, hipMemcpyDeviceToHost));
    return CDFT_OK;
  });
}

int cdft_ddft_open_destroy(cdft_ddft_open d) {
  return guarded([&] {
    delete d;
    return CDFT_OK;
  });
}

// Standalone 3D DST-I of a compact (n0, n1, n2) host array, in place (test and interop helper).
int cdft_dst1_3d(int n0, int n1, int n2, double* a) {
  return guarded([&] {
    REQUIRE(a && n0 > 0 && n1 > 0 && n2 > 0, "bad argument");
    const int ns[3] = {n0, n1, n2};
    for (int ax = 0; ax < 3; ++ax)
      REQUIRE(smooth23(2 * (ns[ax] + 1)) && 2 * (ns[ax] + 1) <= 1024, "DST-I lengths n need 2(n+1) a product of 2 and 3, <= 1024");
    OwnedStream owned;
    hipStream_t st = owned.s;
    const long ni = (long)n0 * n1 * n2;
    DevBuf buf, tw[3];
    buf.alloc(sizeof(double) * ni);
    HIPCHK(hipMemcpyAsync(buf.p, a, sizeof(double) * ni, hipMemcpyHostToDevice, st));
    for (int ax = 2; ax >= 0; --ax) {
      const int L = 2 * (ns[ax] + 1);
      const Radices rad = make_radices(L);
      upload_twiddles<double>(tw[ax], L);
      const int nbl = std::max(1, (256 * kDstPts) / L);
      const long nlines = ni / ns[ax];
      hipLaunchKernelGGL(k_dst1, dim3(div_up(nlines, nbl)), dim3(256), sizeof(cplx<double>) * (size_t)nbl * L, st,
                         buf.as<double>(), n0, n1, n2, ax, rad, tw[ax].as<cplx<double>>(), nbl);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(a, buf.p, sizeof(double) * ni, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return CDFT_OK;
  });
}

}  // extern "C"
