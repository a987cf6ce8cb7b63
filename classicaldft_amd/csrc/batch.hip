// Batched systems, device-resident FIRE2 over a batch, and the tile autotuner.
//
// A batch is one cdft_system with nb members of the same grid and functional
// (shared weight spectra; own rho, dF, V_ext, mu, spectra, partial sums and
// eta flags): every pipeline kernel takes the member from blockIdx.y.
//
// cdft_fire_batch runs fireMinimizer2 (src/Minimizer.cpp:117-327) on every
// member with all scalar control on the device: single-block control kernels
// read the reductions, update dt / alpha / the counters / the back-tracking
// state and the convergence flag, and the vector kernels read those scalars
// from device memory.  One step is therefore a fixed kernel sequence with no
// host round trip, which is captured once into a hipGraph (K steps per graph)
// and replayed; converged or failed members go inactive and stop changing.
#include <map>
#include <mutex>
#include <tuple>

#include "system.hpp"

namespace cdft_impl {
namespace {

constexpr int kBB = 256;  // vector-kernel blocks per member

// Block reduction of NS compensated sums and NM maxima into out[0..NS+NM)
// (sums as (s, c), maxima as (m, 0)); deterministic.
template <int NS, int NM>
__device__ void block_reduce(CSum* s, double* mx, double2* __restrict__ out) {
  __shared__ double ws[NS + NM > 0 ? NS + NM : 1][kThreads / 64][2];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const double s2 = __shfl_xor(s[q].s, off), c2 = __shfl_xor(s[q].c, off);
      s[q].merge(s2, c2);
    }
#pragma unroll
    for (int q = 0; q < NM; ++q) mx[q] = fmax(mx[q], __shfl_xor(mx[q], off));
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      ws[q][wid][0] = s[q].s;
      ws[q][wid][1] = s[q].c;
    }
#pragma unroll
    for (int q = 0; q < NM; ++q) ws[NS + q][wid][0] = mx[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      CSum a{ws[q][0][0], ws[q][0][1]};
      for (int w = 1; w < kThreads / 64; ++w) a.merge(ws[q][w][0], ws[q][w][1]);
      out[q] = make_double2(a.s, a.c);
    }
#pragma unroll
    for (int q = 0; q < NM; ++q) {
      double m = ws[NS + q][0][0];
      for (int w = 1; w < kThreads / 64; ++w) m = fmax(m, ws[NS + q][w][0]);
      out[NS + q] = make_double2(m, 0.0);
    }
  }
}

// Per-member FIRE2 state (fireMinimizer2 members, include/Minimizer.h:105-160).
struct FireCtl {
  double dt, dt_max, dt_best, alpha, vv, F, fold, fmax, fmax_raw, rms_force, vnorm, P_next, P, mix;
  int N_P_positive, N_P_negative, backtracks, it, steps, calls;
  int status;  // 0 running, 1 converged, 2 cannot stop going uphill
  int active, uphill, restore;
};

struct CtlParams {
  cdft_fire_params p;
  Alias al;
  long N;
  double dV, tol;
  // F parts (src/DFT.cpp:377-425 terms present in the functional)
  int has_fmt, ideal, mf, ext;
};

__device__ double member_F(const double* res, double mu, const CtlParams& cp) {
  const double F_hs = cp.has_fmt ? res[0] * cp.dV : 0.0;
  const double F_id = cp.ideal ? res[1] : 0.0;
  const double F_mf = cp.mf ? 0.5 * res[2] : 0.0;
  const double F_ext = (cp.ext ? res[3] * cp.dV : 0.0) - res[4] * cp.dV * mu;
  CSum a{0.0, 0.0};
  a.add(F_id);
  a.add(F_hs);
  a.add(F_mf);
  a.add(F_ext);
  return a.s + a.c;
}

#define MEMBER_LOOP(i) for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < N; i += (long)gridDim.x * kThreads)

// fireMinimizer2::step, the part before SemiImplicitEuler (Minimizer.cpp:160-222).  One thread per member.
__global__ void k_fb_begin(FireCtl* __restrict__ st, int nb, CtlParams cp) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= nb) return;
  FireCtl c = st[y];
  c.uphill = 0;
  c.restore = 0;
  if (c.status != 0) {
    c.active = 0;
    st[y] = c;
    return;
  }
  c.active = 1;
  c.it++;
  c.steps++;
  c.P = c.P_next;
  c.fold = c.F;
  const cdft_fire_params& p = cp.p;
  if (c.P > 0) {
    c.N_P_positive++;
    c.N_P_negative = 0;
    if (c.N_P_positive > p.N_delay) {
      c.dt = fmin(c.dt * p.f_inc, c.dt_max);
      c.alpha = c.alpha * p.f_alf;
    }
  } else {
    c.N_P_positive = 0;
    c.N_P_negative++;
    if (c.N_P_negative > p.N_P_negative_max) {  // "Cannot stop going uphill in Fire2"
      c.status = 2;
      c.active = 0;
      st[y] = c;
      return;
    }
    if (!(p.initial_delay && c.it < (unsigned)p.N_delay)) {
      if (c.dt * p.f_dec >= p.dt_min) c.dt *= p.f_dec;
      c.alpha = p.alpha_start;
    }
    c.uphill = 1;
    c.backtracks++;
  }
  st[y] = c;
}

// Save x, v, dF, dF_raw; on an uphill step x <- x - dt/2 v, v <- v/10.
__global__ void k_fb_save(double* __restrict__ x, double* __restrict__ v, const double* __restrict__ dF,
                          const double* __restrict__ raw, double* __restrict__ xr, double* __restrict__ vr,
                          double* __restrict__ dFr, double* __restrict__ rawr, const FireCtl* __restrict__ st, long N) {
  const FireCtl& c = st[blockIdx.y];
  if (!c.active) return;
  const long o = blockIdx.y * N;
  const double h = -0.5 * c.dt;
  MEMBER_LOOP(i) {
    const double xi = x[o + i], vi = v[o + i];
    xr[o + i] = xi;
    vr[o + i] = vi;
    dFr[o + i] = dF[o + i];
    rawr[o + i] = raw[o + i];
    if (c.uphill) {
      x[o + i] = xi + vi * h;
      v[o + i] = vi * 0.1;
    }
  }
}

// v <- v - dt dF; partials of sum v^2, sum dF^2 (SemiImplicitEuler, Minimizer.cpp:277-292).
__global__ void k_fb_kick(double* __restrict__ v, const double* __restrict__ dF, const FireCtl* __restrict__ st,
                          long N, double2* __restrict__ part) {
  const FireCtl& c = st[blockIdx.y];
  if (!c.active) return;
  const long o = blockIdx.y * N;
  CSum s[2] = {{0, 0}, {0, 0}};
  MEMBER_LOOP(i) {
    const double f = dF[o + i];
    const double nv = v[o + i] + f * (-c.dt);
    v[o + i] = nv;
    s[0].add(nv * nv);
    s[1].add(f * f);
  }
  block_reduce<2, 0>(s, nullptr, part + ((long)blockIdx.y * kBB + blockIdx.x) * 2);
}

// Reduce the kick partials: rms force, |v| and the mixing coefficient.  One block per member.
__global__ void k_fb_mixctl(FireCtl* __restrict__ st, const double2* __restrict__ part, CtlParams cp) {
  FireCtl& c = st[blockIdx.x];
  if (!c.active) return;
  CSum s[2];
  const double2 a = part[((long)blockIdx.x * kBB + threadIdx.x) * 2], b = part[((long)blockIdx.x * kBB + threadIdx.x) * 2 + 1];
  s[0] = {a.x, a.y};
  s[1] = {b.x, b.y};
  __shared__ double2 out[2];
  block_reduce<2, 0>(s, nullptr, out);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double vn2 = out[0].x + out[0].y, fn2 = out[1].x + out[1].y;
    c.rms_force = sqrt(fn2 / cp.N);
    c.vnorm = sqrt(vn2 / cp.N);
    c.mix = fn2 > 0 ? -c.alpha * sqrt(vn2 / fn2) : 0.0;
  }
}

// v <- (1-a) v + mix dF; x <- x + dt v; rho <- alias(x) (Minimizer.cpp:295-300).
__global__ void k_fb_mix(double* __restrict__ x, double* __restrict__ v, const double* __restrict__ dF,
                         double* __restrict__ rho, const FireCtl* __restrict__ st, long N, Alias al) {
  const FireCtl& c = st[blockIdx.y];
  if (!c.active) return;
  const long o = blockIdx.y * N;
  MEMBER_LOOP(i) {
    double nv = v[o + i] * (1 - c.alpha);
    nv += dF[o + i] * c.mix;
    v[o + i] = nv;
    const double nx = x[o + i] + nv * c.dt;
    x[o + i] = nx;
    rho[o + i] = al.rho(nx);
  }
}

// After the evaluation: eta flag -> restore, else F from the member's results.  One thread per member.
__global__ void k_fb_after(FireCtl* __restrict__ st, int* __restrict__ flag, const double* __restrict__ results,
                           const double* __restrict__ mus, int nb, CtlParams cp, int init) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= nb) return;
  const int fl = flag[y];
  if (fl) flag[y] = 0;
  FireCtl& c = st[y];
  if (!c.active) return;
  if (fl) {
    c.restore = 1;
    if (init) c.status = 2;
    return;
  }
  c.restore = 0;
  c.calls++;
  c.F = member_F(results + 8 * y, mus[y], cp);
}

// dF_raw -> alias derivative; partials of sum dF_a^2, sum v.dF_a, max|dF_a|, max|dF_raw|.
__global__ void k_fb_post(const double* __restrict__ x, const double* __restrict__ v, const double* __restrict__ raw,
                          double* __restrict__ dFa, const FireCtl* __restrict__ st, long N, Alias al,
                          double2* __restrict__ part) {
  const FireCtl& c = st[blockIdx.y];
  if (!c.active || c.restore) return;
  const long o = blockIdx.y * N;
  CSum s[2] = {{0, 0}, {0, 0}};
  double mx[2] = {0.0, 0.0};
  MEMBER_LOOP(i) {
    const double r = raw[o + i];
    const double a = r * al.dfac(x[o + i]);
    dFa[o + i] = a;
    s[0].add(a * a);
    s[1].add(v[o + i] * a);
    mx[0] = fmax(mx[0], fabs(a));
    mx[1] = fmax(mx[1], fabs(r));
  }
  block_reduce<2, 2>(s, mx, part + ((long)blockIdx.y * kBB + blockIdx.x) * 4);
}

// End of the step (Minimizer.cpp:225-271, getDF_DX :90-115): back-tracking or the
// post-step bookkeeping, and the convergence test on the raw gradient.  One block per member.
__global__ void k_fb_endctl(FireCtl* __restrict__ st, const double2* __restrict__ part, CtlParams cp, int init) {
  FireCtl& c = st[blockIdx.x];
  if (!c.active) return;
  if (c.restore) {
    if (threadIdx.x == 0) {
      c.dt /= 2;
      c.fmax = 1000;
      c.backtracks++;
      c.P_next = 0.5 * c.P;  // -(v/2).dF of the restored state
      c.F = c.fold;
      if (c.backtracks >= 10) {
        c.dt_max = fmin(c.dt_best, 0.9 * c.dt_max);
        c.dt_best = 0;
        c.backtracks = 0;
      }
    }
    return;
  }
  const double2* pp = part + ((long)blockIdx.x * kBB + threadIdx.x) * 4;
  CSum s[2] = {{pp[0].x, pp[0].y}, {pp[1].x, pp[1].y}};
  double mx[2] = {pp[2].x, pp[3].x};
  __shared__ double2 out[4];
  block_reduce<2, 2>(s, mx, out);
  __syncthreads();
  if (threadIdx.x == 0) {
    c.rms_force = sqrt((out[0].x + out[0].y) / cp.N);
    c.P_next = -(out[1].x + out[1].y);
    c.fmax = out[2].x / cp.dV;
    c.fmax_raw = out[3].x / cp.dV;
    if (!init) {
      c.vv = c.vv * 0.9 + 0.1 * (fabs(c.F - c.fold) / c.dt);
      c.dt_best = fmax(c.dt, c.dt_best);
      if (c.backtracks >= 10) {
        c.dt_max = fmin(c.dt_best, 0.9 * c.dt_max);
        c.dt_best = 0;
        c.backtracks = 0;
      }
    }
    if (c.fmax_raw <= cp.tol && c.fmax >= 0 && c.fmax < 1000) c.status = 1;
  }
}

// Back-tracking restore: x, v/2, dF, dF_raw from the saved copies (Minimizer.cpp:233-245).
__global__ void k_fb_restore(double* __restrict__ x, double* __restrict__ v, double* __restrict__ dF,
                             double* __restrict__ raw, const double* __restrict__ xr, const double* __restrict__ vr,
                             const double* __restrict__ dFr, const double* __restrict__ rawr,
                             const FireCtl* __restrict__ st, long N) {
  const FireCtl& c = st[blockIdx.y];
  if (!c.active || !c.restore) return;
  const long o = blockIdx.y * N;
  MEMBER_LOOP(i) {
    x[o + i] = xr[o + i];
    v[o + i] = vr[o + i] * 0.5;
    dF[o + i] = dFr[o + i];
    raw[o + i] = rawr[o + i];
  }
}

__global__ void k_fb_alias_init(const double* __restrict__ rho0, double* __restrict__ x, double* __restrict__ rho,
                                double* __restrict__ v, long N, Alias al) {
  const long o = blockIdx.y * N;
  MEMBER_LOOP(i) {
    const double xi = al.x_of(rho0[o + i]);
    x[o + i] = xi;
    rho[o + i] = al.rho(xi);
    v[o + i] = 0.0;
  }
}

// Restart of the FIRE2 control for a precision change: the members that met the
// old tolerance run again, time step / mixing / counters start afresh (the
// velocities are zeroed separately); step and call counts are kept.
__global__ void k_fb_reopen(FireCtl* __restrict__ st, int nb, cdft_fire_params p) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= nb) return;
  FireCtl c = st[y];
  if (c.status == 1) c.status = 0;
  c.active = c.status == 0;
  c.restore = 0;
  c.uphill = 0;
  c.dt = p.dt;
  c.dt_max = p.dt_max;
  c.dt_best = p.dt;
  c.alpha = p.alpha_start;
  c.N_P_positive = c.N_P_negative = c.backtracks = 0;
  c.it = 0;
  c.vv = 1.0;
  st[y] = c;
}

__global__ void k_fill(double* __restrict__ p, long n, double v) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads) p[i] = v;
}

__global__ void k_zero_ints(int* __restrict__ p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

#undef MEMBER_LOOP

}  // namespace
}  // namespace cdft_impl

using namespace cdft_impl;

struct cdft_fire_batch_s {
  cdft_system sys = nullptr;        // the system evaluating now (one of sysp)
  cdft_system sysp[2] = {nullptr, nullptr};  // owned, by precision (FP64, MIXED), created on demand
  cdft_system_desc desc{};
  std::vector<double> mus, vexts;  // host copies for creating the other precision
  int nb = 0;
  long N = 0;
  CtlParams cp{};
  DevBuf x, v, dF, raw, rho, xr, vr, dFr, rawr, part, ctl;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
  double graph_tol = -1;

  ~cdft_fire_batch_s() {
    drop_graph();
    for (cdft_system q : sysp)
      if (q) cdft_system_destroy(q);
  }

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }

  dim3 vgrid() const { return dim3(kBB, nb); }

  // Evaluation + post-processing of the current rho (getDF_DX).
  void eval_and_post(int init) {
    hipStream_t st = sys->stream;
    enqueue_eval(sys, rho.as<double>(), raw.as<double>(), true, false);
    hipLaunchKernelGGL(k_fb_after, dim3(div_up(nb, 64)), dim3(64), 0, st, ctl.as<FireCtl>(), sys->flag.as<int>(),
                       sys->results.as<double>(), sys->mus.as<double>(), nb, cp, init);
    hipLaunchKernelGGL(k_fb_post, vgrid(), dim3(kThreads), 0, st, x.as<double>(), v.as<double>(), raw.as<double>(),
                       dF.as<double>(), ctl.as<FireCtl>(), N, cp.al, part.as<double2>());
    hipLaunchKernelGGL(k_fb_endctl, dim3(nb), dim3(kBB), 0, st, ctl.as<FireCtl>(), part.as<double2>(), cp, init);
    HIPCHK(hipGetLastError());
  }

  // One fireMinimizer2::step of every active member, enqueued on the system's stream.  The evaluation kernels skip
  // the members k_fb_begin leaves inactive (FireCtl::active, read on the device by each block).
  void enqueue_step() {
    hipStream_t st = sys->stream;
    FireCtl* c = ctl.as<FireCtl>();
    static_assert(sizeof(FireCtl) % sizeof(int) == 0, "FireCtl is read as an int array");
    sys->skip = Skip{&c->active, long(sizeof(FireCtl) / sizeof(int))};
    struct Reset {
      cdft_system_s* s;
      ~Reset() { s->skip = Skip{}; }
    } reset{sys};
    hipLaunchKernelGGL(k_fb_begin, dim3(div_up(nb, 64)), dim3(64), 0, st, c, nb, cp);
    hipLaunchKernelGGL(k_fb_save, vgrid(), dim3(kThreads), 0, st, x.as<double>(), v.as<double>(), dF.as<double>(),
                       raw.as<double>(), xr.as<double>(), vr.as<double>(), dFr.as<double>(), rawr.as<double>(), c, N);
    hipLaunchKernelGGL(k_fb_kick, vgrid(), dim3(kThreads), 0, st, v.as<double>(), dF.as<double>(), c, N,
                       part.as<double2>());
    hipLaunchKernelGGL(k_fb_mixctl, dim3(nb), dim3(kBB), 0, st, c, part.as<double2>(), cp);
    hipLaunchKernelGGL(k_fb_mix, vgrid(), dim3(kThreads), 0, st, x.as<double>(), v.as<double>(), dF.as<double>(),
                       rho.as<double>(), c, N, cp.al);
    HIPCHK(hipGetLastError());
    eval_and_post(0);
    hipLaunchKernelGGL(k_fb_restore, vgrid(), dim3(kThreads), 0, st, x.as<double>(), v.as<double>(), dF.as<double>(),
                       raw.as<double>(), xr.as<double>(), vr.as<double>(), dFr.as<double>(), rawr.as<double>(), c, N);
    HIPCHK(hipGetLastError());
  }
};

extern "C" {

int cdft_system_batch_create(const cdft_system_desc* desc, int nb, const double* mu, const double* vext,
                             cdft_system* out) {
  return guarded([&] {
    REQUIRE(desc && out && mu && nb >= 1, "bad argument");
    cdft_system s = nullptr;
    create_system(desc, &s, nb, mu, vext);
    *out = s;
    return (int)CDFT_OK;
  });
}

int cdft_fire_batch_create(const cdft_system_desc* desc, int nb, const double* mu, const double* vext,
                           const double* rho0, const cdft_fire_params* params, cdft_fire_batch* out) {
  return guarded([&] {
    REQUIRE(desc && out && mu && rho0 && nb >= 1, "bad argument");
    REQUIRE(!(desc->fixed_mass > 0) && desc->boundary == 0,
            "batched FIRE runs open (fixed-mu) periodic systems: no fixed mass or boundary constraints");
    auto b = std::make_unique<cdft_fire_batch_s>();
    cdft_system s = nullptr;
    create_system(desc, &s, nb, mu, vext);
    b->sys = s;
    b->sysp[desc->precision == CDFT_MIXED ? 1 : 0] = s;
    b->desc = *desc;
    b->desc.vext = nullptr;
    b->mus.assign(mu, mu + nb);
    if (vext) b->vexts.assign(vext, vext + (size_t)s->N * nb);
    b->nb = nb;
    b->N = s->N;
    const long n = s->N * nb;
    for (DevBuf* q : {&b->x, &b->v, &b->dF, &b->raw, &b->rho, &b->xr, &b->vr, &b->dFr, &b->rawr})
      q->alloc(sizeof(double) * n);
    b->part.alloc(sizeof(double2) * 4 * kBB * nb);
    b->ctl.alloc(sizeof(FireCtl) * nb);
    CtlParams& cp = b->cp;
    if (params) cp.p = *params;
    else cdft_fire_default_params(&cp.p);
    cp.N = s->N;
    cp.dV = s->d.dx * s->d.dy * s->d.dz;
    cp.tol = 0.0;
    cp.al.fmt = s->m > 0;
    const double d = s->d.hsd;
    cp.al.c = (1.0 - 1e-18 * (4 * M_PI * d * d * d / 3)) / cp.dV;
    cp.has_fmt = s->m > 0;
    cp.ideal = s->d.ideal;
    cp.mf = s->d.mean_field;
    cp.ext = s->has_vext;
    std::vector<FireCtl> c0(nb);
    for (auto& c : c0) {
      c = FireCtl{};
      c.dt = cp.p.dt;
      c.dt_max = cp.p.dt_max;
      c.dt_best = cp.p.dt;
      c.alpha = cp.p.alpha_start;
      c.vv = 1.0;
      c.fmax = -1;
      c.active = 1;
    }
    HIPCHK(hipMemcpy(b->ctl.p, c0.data(), sizeof(FireCtl) * nb, hipMemcpyHostToDevice));
    DevBuf r0;
    r0.alloc(sizeof(double) * n);
    HIPCHK(hipMemcpy(r0.p, rho0, sizeof(double) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_fb_alias_init, b->vgrid(), dim3(kThreads), 0, s->stream, r0.as<double>(), b->x.as<double>(),
                       b->rho.as<double>(), b->v.as<double>(), b->N, cp.al);
    HIPCHK(hipGetLastError());
    b->eval_and_post(1);
    HIPCHK(hipStreamSynchronize(s->stream));
    *out = b.release();
    return (int)CDFT_OK;
  });
}

int cdft_fire_batch_run(cdft_fire_batch b, int steps, double tol_raw, int use_graph) {
  return guarded([&] {
    REQUIRE(b && steps > 0, "bad argument");
    b->cp.tol = tol_raw;
    hipStream_t st = b->sys->stream;
    if (!use_graph) {
      for (int k = 0; k < steps; ++k) b->enqueue_step();
      return (int)CDFT_OK;
    }
    if (!b->exec || b->graph_steps != steps || b->graph_tol != tol_raw) {
      b->drop_graph();
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        for (int k = 0; k < steps; ++k) b->enqueue_step();
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIPCHK(hipStreamEndCapture(st, &b->graph));
      HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_steps = steps;
      b->graph_tol = tol_raw;
    }
    HIPCHK(hipGraphLaunch(b->exec, st));
    return (int)CDFT_OK;
  });
}

int cdft_fire_batch_set_precision(cdft_fire_batch b, int precision) {
  return guarded([&] {
    REQUIRE(b && (precision == CDFT_FP64 || precision == CDFT_MIXED), "bad argument");
    const int k = precision == CDFT_MIXED ? 1 : 0;
    if (b->sys == b->sysp[k]) return (int)CDFT_OK;
    HIPCHK(hipStreamSynchronize(b->sys->stream));
    if (!b->sysp[k]) {
      cdft_system_desc d = b->desc;
      d.precision = precision;
      create_system(&d, &b->sysp[k], b->nb, b->mus.data(), b->vexts.empty() ? nullptr : b->vexts.data());
    }
    b->sys = b->sysp[k];
    b->drop_graph();
    // members that met the tolerance at the old precision run again; the
    // current densities are re-evaluated at the new one
    hipLaunchKernelGGL(k_fb_reopen, dim3(div_up(b->nb, 64)), dim3(64), 0, b->sys->stream, b->ctl.as<FireCtl>(), b->nb,
                       b->cp.p);
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(kThreads), 0, b->sys->stream, b->v.as<double>(), b->N * b->nb, 0.0);
    HIPCHK(hipGetLastError());
    b->cp.tol = -1.0;  // no convergence test on the re-evaluation; the next run sets the tolerance
    b->eval_and_post(1);
    HIPCHK(hipStreamSynchronize(b->sys->stream));
    return (int)CDFT_OK;
  });
}

int cdft_fire_batch_status(cdft_fire_batch b, cdft_fire_report* reports, int* status, int* running) {
  return guarded([&] {
    REQUIRE(b, "null argument");
    std::vector<FireCtl> c(b->nb);
    HIPCHK(hipMemcpyAsync(c.data(), b->ctl.p, sizeof(FireCtl) * b->nb, hipMemcpyDeviceToHost, b->sys->stream));
    HIPCHK(hipStreamSynchronize(b->sys->stream));
    int run = 0;
    for (int i = 0; i < b->nb; ++i) {
      run += c[i].status == 0;
      if (status) status[i] = c[i].status;
      if (reports) {
        cdft_fire_report& r = reports[i];
        r.F = c[i].F;
        r.fmax_alias = c[i].fmax;
        r.fmax_raw = c[i].fmax_raw;
        r.dt = c[i].dt;
        r.alpha = c[i].alpha;
        r.vv = c[i].vv;
        r.rms_force = c[i].rms_force;
        r.steps = c[i].steps;
        r.calls = c[i].calls;
        r.monitor = std::fabs(c[i].vv / (b->N * b->cp.dV));
      }
    }
    if (running) *running = run;
    return (int)CDFT_OK;
  });
}

int cdft_fire_batch_get(cdft_fire_batch b, int member, double* rho) {
  return guarded([&] {
    REQUIRE(b && rho && member >= 0 && member < b->nb, "bad argument");
    HIPCHK(hipMemcpyAsync(rho, b->rho.as<double>() + member * b->N, sizeof(double) * b->N, hipMemcpyDeviceToHost,
                          b->sys->stream));
    HIPCHK(hipStreamSynchronize(b->sys->stream));
    return (int)CDFT_OK;
  });
}

int cdft_fire_batch_state(cdft_fire_batch b, int member, double* rho, double* dF_raw, double* x, double* v) {
  return guarded([&] {
    REQUIRE(b && member >= 0 && member < b->nb, "bad argument");
    const long o = member * b->N;
    const size_t bytes = sizeof(double) * b->N;
    if (rho) HIPCHK(hipMemcpyAsync(rho, b->rho.as<double>() + o, bytes, hipMemcpyDeviceToHost, b->sys->stream));
    if (dF_raw) HIPCHK(hipMemcpyAsync(dF_raw, b->raw.as<double>() + o, bytes, hipMemcpyDeviceToHost, b->sys->stream));
    if (x) HIPCHK(hipMemcpyAsync(x, b->x.as<double>() + o, bytes, hipMemcpyDeviceToHost, b->sys->stream));
    if (v) HIPCHK(hipMemcpyAsync(v, b->v.as<double>() + o, bytes, hipMemcpyDeviceToHost, b->sys->stream));
    HIPCHK(hipStreamSynchronize(b->sys->stream));
    return (int)CDFT_OK;
  });
}

int cdft_fire_batch_system(cdft_fire_batch b, cdft_system* sys) {
  return guarded([&] {
    REQUIRE(b && sys, "null argument");
    *sys = b->sys;
    return (int)CDFT_OK;
  });
}

void cdft_fire_batch_destroy(cdft_fire_batch b) { delete b; }

}  // extern "C"

// A group of FIRE batches (typically one per grid size) stepped together: K steps of every batch are captured into
// one hipGraph, each batch on its own stream forked from the group's stream and joined back, so one graph launch
// advances systems of different grids concurrently.  The graph is re-captured when the step count, the tolerance
// or a batch's precision changes.
struct cdft_fire_group_s {
  std::vector<cdft_fire_batch> b;
  std::vector<cdft_system> sys_at_capture;
  hipStream_t st = nullptr;
  hipEvent_t fork = nullptr, done = nullptr;
  std::vector<hipEvent_t> join, pre;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int steps = 0;
  double tol = -1;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~cdft_fire_group_s() {
    drop_graph();
    for (hipEvent_t e : join) (void)hipEventDestroy(e);
    for (hipEvent_t e : pre) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (done) (void)hipEventDestroy(done);
    if (st) (void)hipStreamDestroy(st);
  }
};

extern "C" {

int cdft_fire_group_create(cdft_fire_batch* batches, int n, cdft_fire_group* out) {
  return guarded([&] {
    REQUIRE(batches && out && n >= 1, "bad argument");
    auto g = std::make_unique<cdft_fire_group_s>();
    for (int i = 0; i < n; ++i) {
      REQUIRE(batches[i], "null batch");
      for (int j = 0; j < i; ++j) REQUIRE(batches[j] != batches[i], "a batch appears twice in the group");
      g->b.push_back(batches[i]);
    }
    HIPCHK(hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&g->done, hipEventDisableTiming));
    g->join.resize(n);
    g->pre.resize(n);
    for (int i = 0; i < n; ++i) {
      HIPCHK(hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&g->pre[i], hipEventDisableTiming));
    }
    *out = g.release();
    return (int)CDFT_OK;
  });
}

int cdft_fire_group_run(cdft_fire_group g, int steps, double tol_raw) {
  return guarded([&] {
    REQUIRE(g && steps > 0, "bad argument");
    const int n = (int)g->b.size();
    bool stale = !g->exec || g->steps != steps || g->tol != tol_raw;
    for (int i = 0; i < n && !stale; ++i) stale = g->sys_at_capture[i] != g->b[i]->sys;
    for (cdft_fire_batch b : g->b) b->cp.tol = tol_raw;
    if (stale) {
      g->drop_graph();
      HIPCHK(hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal));
      try {
        HIPCHK(hipEventRecord(g->fork, g->st));
        for (int i = 0; i < n; ++i) {
          hipStream_t bs = g->b[i]->sys->stream;
          HIPCHK(hipStreamWaitEvent(bs, g->fork, 0));
          for (int k = 0; k < steps; ++k) g->b[i]->enqueue_step();
          HIPCHK(hipEventRecord(g->join[i], bs));
          HIPCHK(hipStreamWaitEvent(g->st, g->join[i], 0));
        }
      } catch (...) {
        hipGraph_t gr = nullptr;
        (void)hipStreamEndCapture(g->st, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
      }
      HIPCHK(hipStreamEndCapture(g->st, &g->graph));
      HIPCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
      g->steps = steps;
      g->tol = tol_raw;
      g->sys_at_capture.clear();
      for (cdft_fire_batch b : g->b) g->sys_at_capture.push_back(b->sys);
    }
    // the launch follows the work already queued on each batch's stream, and that stream follows the launch
    for (int i = 0; i < n; ++i) {
      HIPCHK(hipEventRecord(g->pre[i], g->b[i]->sys->stream));
      HIPCHK(hipStreamWaitEvent(g->st, g->pre[i], 0));
    }
    HIPCHK(hipGraphLaunch(g->exec, g->st));
    HIPCHK(hipEventRecord(g->done, g->st));
    for (cdft_fire_batch b : g->b) HIPCHK(hipStreamWaitEvent(b->sys->stream, g->done, 0));
    return (int)CDFT_OK;
  });
}

void cdft_fire_group_destroy(cdft_fire_group g) {
  if (g) (void)hipStreamSynchronize(g->st);
  delete g;
}

int cdft_system_autotune(cdft_system s, int reps, double* best_ms) {
  return guarded([&] {
    REQUIRE(s && reps > 0, "bad argument");
    struct Cfg {
      bool reg, cols, fwd, bwd;
      int pw;
      long slab;
    };
    struct Key {
      int Nx, Ny, Nz, mixed, model, nconv, nb;
      bool operator<(const Key& o) const {
        return std::tie(Nx, Ny, Nz, mixed, model, nconv, nb) < std::tie(o.Nx, o.Ny, o.Nz, o.mixed, o.model, o.nconv, o.nb);
      }
    };
    static std::map<Key, std::pair<Cfg, double>> cache;
    static std::mutex mtx;
    const Key key{s->g.Nx, s->g.Ny, s->g.Nz, (int)s->mixed, s->d.model, s->nconv, s->nb};
    auto apply = [&](const Cfg& c) {
      s->reg_fft = c.reg || s->wq;  // folded weight spectra are only read by the register x pass
      s->nb_half_cols = c.cols;
      s->nb_half_fwd = c.fwd;
      s->nb_half_bwd = c.bwd;
      s->pw_rows = c.pw;
      s->slab_bytes = c.slab;
    };
    {
      std::lock_guard<std::mutex> lk(mtx);
      auto it = cache.find(key);
      if (it != cache.end()) {
        apply(it->second.first);
        if (best_ms) *best_ms = it->second.second;
        return (int)CDFT_OK;
      }
    }
    std::vector<Cfg> cands;
    const bool reg_possible = reg_fft_ok<double>(*s, s->g.Nx) || rows_reg_ok<double>(*s);
    // batches and folded weight spectra (ColGeo::wq) run the register x pass only
    for (int r = reg_possible ? 1 : 0; r >= ((s->nb > 1 || s->wq) ? 1 : 0); --r)
      for (int cols = 0; cols < 2; ++cols)
        for (int fwd = 0; fwd < 2; ++fwd)
          for (int bwd = 0; bwd < 2; ++bwd)
            for (int pw : {0, 1}) {
              if (!r && (cols || fwd || bwd || pw)) continue;  // the LDS kernels have no such choices
              cands.push_back({r == 1, cols == 1, fwd == 1, bwd == 1, pw, s->slab_bytes});
            }
    // a uniform density in the member buffers (no eta flags, realistic work)
    const long n = s->N * s->nb;
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(kThreads), 0, s->stream, s->rho.as<double>(), n, 0.05);
    HIPCHK(hipGetLastError());
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e300;
    Cfg bestc = cands.front();
    auto trial = [&](const Cfg& c) {
      apply(c);
      enqueue_eval(s, s->rho.as<double>(), s->dF.as<double>(), true, false);  // warm-up
      HIPCHK(hipEventRecord(e0, s->stream));
      for (int i = 0; i < reps; ++i) enqueue_eval(s, s->rho.as<double>(), s->dF.as<double>(), true, false);
      HIPCHK(hipEventRecord(e1, s->stream));
      HIPCHK(hipEventSynchronize(e1));
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, e0, e1));
      if (t / reps < best) {
        best = t / reps;
        bestc = c;
      }
    };
    for (const Cfg& c : cands) trial(c);
    // then the x-slab size of the y / z / y section on the best tiles (0: three whole-array passes)
    if (bestc.reg)
      for (long mb : {0L, 96L, 192L, 288L}) {
        Cfg c = bestc;
        c.slab = mb << 20;
        if (c.slab != bestc.slab) trial(c);
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(kThreads), 0, s->stream, s->flag.as<int>(), s->nb);
    HIPCHK(hipStreamSynchronize(s->stream));
    apply(bestc);
    {
      std::lock_guard<std::mutex> lk(mtx);
      cache[key] = {bestc, best};
    }
    if (best_ms) *best_ms = best;
    return (int)CDFT_OK;
  });
}

}  // extern "C"
