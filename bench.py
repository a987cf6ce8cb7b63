#!/usr/bin/env python3
"""Benchmark of the classicalDFT hot path on one MI355X.

Workload (the metric's "@256^3" case): one FMT free energy + dF/drho
evaluation (FMT::calculateFreeEnergyAndDerivatives, src/FMT.cpp:363-427) on a
periodic 256^3 grid, dx = 0.04, hard-sphere diameter 1, White Bear I, density
rho_i = 0.6 (1 + 0.05 U[-1,1]) with U from splitmix64(0x5EED0256).  A "step"
is one evaluation with rho already resident in HBM.  The FP64 path is the
headline value; the mixed path (FP32 FFTs and convolutions, FP64 pointwise
Phi and reductions) is reported beside it with its error against FP64.

The roofline uses the algorithmic byte model of one evaluation,
B(e, m) = e N (4m + 2) + e Nk (2m + 8) (each 3D FFT one read + one write,
every stage array written and read once), N = 256^3, Nk = 256*256*129,
m = 11; the dominant kernel's own compulsory bytes (one pass over the arrays it
touches) and its rate are reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
from classicaldft_amd.workloads import splitmix64_uniform, well_systems  # noqa: E402  (no library load)

SEED = 0x5EED0256
SHAPE = (256, 256, 256)
DX = 0.04
HSD = 1.0
PEAK_HBM_GBS = 8000.0


def workload_density(shape=SHAPE):
    u = splitmix64_uniform(int(np.prod(shape)), SEED).reshape(shape)
    return 0.6 * (1.0 + 0.05 * (2.0 * u - 1.0))


def model_bytes(e, m, shape=SHAPE, fp64_io=False):
    """B(e, m); with fp64_io the density read and dF write stay 8-byte (mixed path)."""
    Nx, Ny, Nz = shape
    N = Nx * Ny * Nz
    Nk = Nx * Ny * (Nz // 2 + 1)
    if fp64_io:
        return e * N * (4 * m) + e * Nk * (2 * m + 8) + 16 * N
    return e * N * (4 * m + 2) + e * Nk * (2 * m + 8)


def stage_bytes(shape, e, m, nconv=None, folded=True):
    """Compulsory HBM bytes of each kernel of this implementation (one pass each over the arrays it touches):
    m stored weighted-density arrays, nconv convolved (m plus a mean-field kernel); weight spectra real-valued,
    e bytes per k-point, rows kx <= Nx/2 only when folded."""
    Nx, Ny, Nz = shape
    nconv = m if nconv is None else nconv
    N = Nx * Ny * Nz
    Nk = Nx * Ny * (Nz // 2 + 1)
    c = 2 * e
    wk = e * Nk * ((Nx // 2 + 1) / Nx if folded else 1.0)  # one weight spectrum
    return {
        "z_r2c": 8 * N + c * Nk,
        "y_fwd": 2 * c * Nk,
        "x_conv_fwd": c * Nk + nconv * wk + nconv * c * Nk,
        "y_inv": 2 * nconv * c * Nk,
        "z_pointwise": nconv * c * Nk + m * c * Nk,
        "y_fwd2": 2 * m * c * Nk,
        "x_conv_bwd": m * c * Nk + m * wk + c * Nk,
        "y_inv2": 2 * c * Nk,
        "z_final": c * Nk + 8 * N + 8 * N,  # acc in, rho in, dF out
    }


def time_system(cdft, precision, rho_d, dF_d, steps, warmup):
    s = cdft.System(SHAPE, DX, HSD, model=cdft.WHITEBEAR_I, precision=precision)
    layout = s.layout
    for _ in range(warmup):
        s.eval_async(rho_d.ptr, dF_d.ptr, fmt_only=True)
    parts = s.fetch()
    ms = s.time(rho_d.ptr, dF_d.ptr, steps, fmt_only=True)
    parts = s.fetch()
    # per-stage durations (hipEvents between the kernels of one evaluation)
    s.set_timing(True)
    acc = {}
    reps = 3
    for _ in range(reps):
        s.eval_async(rho_d.ptr, dF_d.ptr, fmt_only=True)
        s.fetch()
        for name, t in s.timing():
            acc[name] = acc.get(name, 0.0) + t / reps
    s.set_timing(False)
    dF = dF_d.download()
    s.close()
    return ms / steps, parts[2], dF, acc, layout


def time_ao(cdft, rho_d, dF_d, steps, Rp=0.25, rho_p=0.5):
    """The same 256^3 evaluation with an Asakura-Oosawa polymer (FMT_AO_Species) of radius Rp."""
    s = cdft.System(SHAPE, DX, HSD, model=cdft.WHITEBEAR_I, ao_Rp=Rp, ao_reservoir=rho_p)
    s.eval_async(rho_d.ptr, dF_d.ptr, fmt_only=True)
    s.fetch()
    ms = s.time(rho_d.ptr, dF_d.ptr, steps, fmt_only=True)
    s.fetch()
    s.close()
    return {"Rp": Rp, "reservoir_density": rho_p, "ms_per_step": ms / steps, "evals_per_s": 1e3 * steps / ms}


def time_mixture(cdft, rho_d, steps, d2=0.6, frac=0.5):
    """A binary hard-sphere mixture on the 256^3 lattice (FMT over two FMT_Species, diameters HSD and d2): the
    workload density split between the species, F+dF of both per evaluation, device pointers, FP64.  Wall clock
    around `steps` evaluations (each returns after its stream has finished)."""
    import time
    mix = cdft.Mixture(SHAPE, DX, [HSD, d2], model=cdft.WHITEBEAR_I)
    rho = rho_d.download()
    r = [cdft.DeviceArray.from_host(frac * rho), cdft.DeviceArray.from_host((1 - frac) * rho)]
    f = [cdft.DeviceArray(SHAPE), cdft.DeviceArray(SHAPE)]
    rp, fp = [a.ptr for a in r], [a.ptr for a in f]
    for _ in range(2):
        mix.fmt_eval_device(rp, fp)
    t0 = time.perf_counter()
    for _ in range(steps):
        F = mix.fmt_eval_device(rp, fp)
    sec = time.perf_counter() - t0
    mix.close()
    return {"species": 2, "hsd": [HSD, d2], "split": [frac, 1 - frac], "ms_per_step": 1e3 * sec / steps,
            "evals_per_s": steps / sec, "F": F, "timing": "wall clock, synchronised evaluations"}


def slit_pore(cdft, n=128, dx=0.05, kT=0.8, rho_b=0.7, precision=0):
    """Slit pore: White Bear I + LJ (sigma = eps = 1, rc = 2.5, WCA, QF kernel) at kT, hard walls
    V_ext(z) = 10 [exp(-z/0.2) + exp(-(L-z)/0.2)], mu = mu_times_beta(rho_b), start rho_b exp(-V_ext)."""
    shape = (n, n, n)
    L = n * dx
    hsd = cdft.lj_hsd(kT)
    z = np.arange(n) * dx
    vz = 10.0 * (np.exp(-z / 0.2) + np.exp(-(L - z) / 0.2))
    vext = np.ascontiguousarray(np.broadcast_to(vz, shape))
    probe = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=kT)
    _, mu = probe.bulk(rho_b)
    probe.close()
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, ideal=True, mu=mu, vext=vext, mean_field=True, kT=kT,
                    precision=precision)
    return s, rho_b * np.exp(-vext), {"grid": list(shape), "dx": dx, "kT": kT, "hsd": hsd, "mu": mu,
                                       "rho_bulk": rho_b}


def fire_time_to_tol(cdft, n=128, tol=1e-8, max_steps=20000, precision=0):
    """FIRE2 from rho_b exp(-V_ext) until max_i |dF/drho_i| / dV <= tol; wall time of the whole run."""
    s, rho0, info = slit_pore(cdft, n=n, precision=precision)
    fire = cdft.Fire(s, rho0)
    t0 = time.perf_counter()
    rc, rep = fire.run(max_steps, tol)
    dt = time.perf_counter() - t0
    info.update({"seconds": dt, "steps": rep["steps"], "evaluations": rep["calls"], "F": rep["F"],
                 "fmax_raw": rep["fmax_raw"], "converged": bool(rep["fmax_raw"] <= tol), "status": rc,
                 "ms_per_step": 1e3 * dt / max(1, rep["steps"])})
    fire.close()
    s.close()
    return info


def fire_ladder_time_to_tol(cdft, n=128, tol=1e-8, ladder_tol=1e-5, max_steps=20000):
    """The same slit pore by the precision ladder: mixed-precision FIRE2 until max |dF/drho| / dV <= ladder_tol
    (the FP32 spectra's gradient floor is ~1e-6), then FP64 FIRE2 from that density to tol; wall time of both."""
    s_mx, rho0, info = slit_pore(cdft, n=n, precision=cdft.MIXED)
    s64, _, _ = slit_pore(cdft, n=n, precision=cdft.FP64)
    t0 = time.perf_counter()
    fm = cdft.Fire(s_mx, rho0)
    _, repm = fm.run(max_steps, ladder_tol)
    rm = fm.state()[0]
    t1 = time.perf_counter()
    f64 = cdft.Fire(s64, rm)
    rc, rep = f64.run(max_steps, tol)
    t2 = time.perf_counter()
    info.update({"seconds": t2 - t0, "seconds_mixed": t1 - t0, "seconds_fp64": t2 - t1, "ladder_tol": ladder_tol,
                 "steps_mixed": repm["steps"], "steps_fp64": rep["steps"], "F": rep["F"], "fmax_raw": rep["fmax_raw"],
                 "converged": bool(rep["fmax_raw"] <= tol), "status": rc})
    for o in (fm, f64, s_mx, s64):
        o.close()
    return info


def fire_graph_time_to_tol(cdft, n=128, tol=1e-8, ladder_tol=None, rounds=25, max_steps=20000):
    """The same slit pore by FIRE2 with device-side control (a FireBatch of one system): `rounds` steps per hipGraph
    launch, the host checking the status between launches; with ladder_tol, mixed precision to ladder_tol first,
    then FP64 to tol.  Wall time of the whole run (batch creation and its first evaluation included)."""
    s, rho0, info = slit_pore(cdft, n=n)
    s.close()
    z = np.arange(n) * info["dx"]
    vz = 10.0 * (np.exp(-z / 0.2) + np.exp(-(n * info["dx"] - z) / 0.2))  # the walls of slit_pore
    vext = np.ascontiguousarray(np.broadcast_to(vz, (n, n, n)))
    t0 = time.perf_counter()
    fb = cdft.FireBatch(tuple(info["grid"]), info["dx"], [info["mu"]], rho0[None], vexts=vext[None],
                        hsd=info["hsd"], precision=cdft.MIXED if ladder_tol else cdft.FP64, ideal=True,
                        mean_field=True, kT=info["kT"])
    phases = ([(cdft.MIXED, ladder_tol)] if ladder_tol else []) + [(cdft.FP64, tol)]
    split = {}
    for prec, t in phases:
        fb.set_precision(prec)
        while True:
            fb.run(rounds, t)
            reps, st, running = fb.status()
            if running == 0 or reps[0]["steps"] >= max_steps:
                break
        split["mixed" if prec == cdft.MIXED else "fp64"] = (time.perf_counter() - t0, reps[0]["steps"])
    dt = time.perf_counter() - t0
    rep = reps[0]
    fb.close()
    info.update({"seconds": dt, "steps": rep["steps"], "F": rep["F"], "fmax_raw": rep["fmax_raw"],
                 "converged": bool(rep["fmax_raw"] <= tol), "status": int(st[0]), "steps_per_launch": rounds,
                 "control": "device (FireBatch of one system, hipGraph of the steps)"})
    if ladder_tol:
        info.update({"ladder_tol": ladder_tol, "seconds_mixed": split["mixed"][0], "steps_mixed": split["mixed"][1]})
    return info


def ddft_fcc_seed(cdft, steps, n=256, dx=0.04, kT=0.8, rho_l=0.94, a=1.55, alpha=50.0, dt=1e-4, precision=0):
    """DDFT from a liquid with a 4x4x4 fcc seed of Gaussians (alpha/pi)^{3/2} exp(-alpha |r - R|^2) centred in the
    box, dt pinned at 1e-4; ms per step and the relative mass drift."""
    shape = (n, n, n)
    L = n * dx
    hsd = cdft.lj_hsd(kT)
    x = np.arange(n) * dx
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.array([[i, j, k] for i in range(4) for j in range(4) for k in range(4)], dtype=float)
    sites = ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a) + (L - 4 * a) / 2
    rho = np.full(shape, rho_l)
    lo, hi = (L - 4 * a) / 2, (L + 4 * a) / 2
    inside = ((x >= lo) & (x < hi))
    sub = np.zeros((inside.sum(),) * 3)
    xi = x[inside]
    gx = np.exp(-alpha * (xi[None, :] - sites[:, 0:1]) ** 2)
    gy = np.exp(-alpha * (xi[None, :] - sites[:, 1:2]) ** 2)
    gz = np.exp(-alpha * (xi[None, :] - sites[:, 2:3]) ** 2)
    sub = np.einsum("si,sj,sk->ijk", gx, gy, gz) * (alpha / np.pi) ** 1.5
    idx = np.where(inside)[0]
    rho[np.ix_(idx, idx, idx)] = np.maximum(sub, 1e-12)
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=kT, precision=precision)
    d = cdft.DDFT(s, rho, dt=dt, dt_max=dt)
    d.set_max_time_step(dt)
    m0 = rho.sum() * dx ** 3
    t0 = time.perf_counter()
    r = None
    done, stalled = 0, None
    for i in range(steps):
        try:
            r = d.step()
        except cdft.cdft.DdftStalled as e:  # the reference's restart loop would not end here
            stalled = str(e)
            break
        done += 1
        if (i + 1) % 100 == 0:
            print(f"# ddft step {i + 1}: dt {r['dt']:.3g}, {time.perf_counter() - t0:.1f} s", flush=True)
    sec = time.perf_counter() - t0
    out = {"grid": list(shape), "dx": dx, "kT": kT, "steps": done, "steps_requested": steps, "seconds": sec,
           "ms_per_step": 1e3 * sec / max(1, done), "F": r["F"] if r else None, "evaluations": r["calls"] if r else 0,
           "mass_rel_drift": abs(r["mass"] - m0) / m0 if r else None, "dt": r["dt"] if r else None,
           "stalled": stalled}
    d.close()
    s.close()
    return out


def batch_fire(cdft, classes=((48, 32), (64, 32), (96, 32)), dx=0.1, kT=0.8, tol=1e-8, ladder_tol=1e-5,
               ladder_rounds=24, check=25, max_rounds=2000, sequential=True):
    """Batched FIRE2 to max |dF/drho| <= tol on systems of several grid sizes: a device-controlled FIRE batch per grid
    class (all members in every kernel), all classes captured into one hipGraph (one launch per round of `check`
    steps, the classes concurrent), tiles autotuned per class, mixed precision until ladder_tol (at most
    ladder_rounds x check steps), then FP64 from a FIRE restart."""
    hsd = cdft.lj_hsd(kT)
    kw = dict(hsd=hsd, model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=kT)
    seed = 0x5EED4000
    batches, info = [], {"classes": [], "tol": tol, "ladder_tol": ladder_tol, "hsd": hsd, "dx": dx, "kT": kT}
    for n, count in classes:
        rho_b, vexts = well_systems(n, count, seed, dx)
        seed += count
        probe = cdft.System((n, n, n), dx, **kw)
        mus = np.array([probe.bulk(r)[1] for r in rho_b])
        probe.close()
        rho0 = np.repeat(rho_b, n ** 3).reshape(count, n, n, n)
        fb = cdft.FireBatch((n, n, n), dx, mus, rho0, vexts=vexts, precision=cdft.MIXED, **kw)
        t_mixed = fb.autotune(reps=3)
        batches.append(fb)
        info["classes"].append({"grid": n, "systems": count, "ms_per_batched_eval_mixed": t_mixed})
    t0 = time.perf_counter()
    phase_t = []
    group = cdft.FireGroup(batches)  # every class advances in one graph launch per round
    for phase, (prec, ptol) in enumerate(((cdft.MIXED, ladder_tol), (cdft.FP64, tol))):
        if phase == 1:
            for b, fb in enumerate(batches):
                fb.set_precision(prec)
                info["classes"][b]["ms_per_batched_eval_fp64"] = fb.autotune(reps=3)
        tp = time.perf_counter()
        last = tp
        for rnd in range(ladder_rounds if phase == 0 else max_rounds):
            group.run(check, ptol)
            worst, nrun = 0.0, 0
            for fb in batches:
                reps, st, running = fb.status()
                if running:
                    nrun += running
                    worst = max(worst, max(r["fmax_raw"] for r, q in zip(reps, st) if q == 0))
            if nrun == 0:
                break
            if time.perf_counter() - last > 10:
                last = time.perf_counter()
                print(f"#   phase {phase} round {rnd}: {nrun} running, worst max|dF/drho| {worst:.3e}, "
                      f"{last - tp:.1f} s", flush=True)
        phase_t.append(time.perf_counter() - tp)
        print(f"#   phase {phase} done in {phase_t[-1]:.2f} s", flush=True)
    total = time.perf_counter() - t0
    group.close()
    steps = conv = fails = 0
    for b, fb in enumerate(batches):
        reps, st, _ = fb.status()
        cs = [r["steps"] for r in reps]
        info["classes"][b].update({"converged": int((st == 1).sum()), "max_steps": int(max(cs)),
                                   "mean_steps": float(np.mean(cs)), "max_fmax_raw": float(max(r["fmax_raw"] for r in reps))})
        steps += sum(cs)
        conv += int((st == 1).sum())
        fails += int((st == 2).sum())
        fb.close()
    info.update({"seconds": total, "seconds_mixed_phase": phase_t[0], "seconds_fp64_phase": phase_t[1],
                 "systems": sum(c for _, c in classes), "converged": conv, "failed": fails, "system_steps": steps,
                 "systems_per_s": conv / total, "system_steps_per_s": steps / total, "check_every": check})
    if sequential:
        # one system per class with the host-controlled single-system FIRE2 (FP64), scaled by the class size
        est = 0.0
        seed = 0x5EED4000
        for n, count in classes:
            rho_b, vexts = well_systems(n, 1, seed, dx)
            seed += count
            s = cdft.System((n, n, n), dx, **kw)
            mu = s.bulk(rho_b[0])[1]
            s.close()
            s = cdft.System((n, n, n), dx, mu=mu, vext=vexts[0], **kw)
            f = cdft.Fire(s, np.full((n, n, n), rho_b[0]))
            t1 = time.perf_counter()
            f.run(100000, tol)
            est += (time.perf_counter() - t1) * count
            f.close()
            s.close()
        info["sequential_single_system_estimate_s"] = est
        info["batched_speedup_vs_sequential"] = est / total
    return info


def host_info():
    """The host CPU the baseline ran on: logical CPUs, lscpu's model name and core layout."""
    info = {"nproc": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys and keys[k.strip()] not in info:
                info[keys[k.strip()]] = v.strip()
    except Exception as e:  # lscpu missing: fall back to /proc/cpuinfo
        info["lscpu_error"] = str(e)
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        info["model"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    try:
        info["physical_cores"] = int(info["sockets"]) * int(info["cores_per_socket"])
    except (KeyError, ValueError):
        pass
    return info


def cpu_fmt(rho, shape, dx, threads, fft_threads, reps):
    """Median seconds of `reps` oracle FMT F+dF evaluations after one warm-up (weight set-up excluded)."""
    import oracle
    oracle.set_threads(threads, fft_threads)
    h = oracle.FMT(oracle.WBI, shape, dx, HSD, weights=oracle.fmt_weights(shape, dx, HSD,
                                                                          mode=oracle.WEIGHTS_ACCURATE))
    F, dF = h.eval(rho)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        F, dF = h.eval(rho)
        ts.append(time.perf_counter() - t0)
    del h
    return float(np.median(ts)), ts, F, dF


_PROBE = r"""
import json, torch
n = (1 << 30) // 8
a = torch.ones(n, dtype=torch.float64, device="cuda")
b = torch.empty_like(a)
out = {"bytes": 8 * n, "tool": "torch copy_ / fill_ / sum, 1 GiB float64, 10 timed after 3 untimed"}
for name, fn, moved in (("copy", lambda: b.copy_(a), 16 * n), ("write", lambda: b.fill_(1.5), 8 * n),
                        ("read", lambda: a.sum(), 8 * n)):
    for _ in range(3):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        fn()
    e1.record()
    torch.cuda.synchronize()
    out[name + "_GBs"] = moved / (e0.elapsed_time(e1) / 10 * 1e-3) / 1e9
print(json.dumps(out))
"""


def streaming_probe():
    """Achievable HBM streaming rates of this GPU, beside the published 8 TB/s peak: a device-to-device copy (read
    + write), a fill (write only) and a sum (read only) of 1 GiB through torch, timed with events on its stream.  A
    child process runs it: torch brings its own HIP runtime, which cannot share a process with the library's."""
    import subprocess
    r = subprocess.run([sys.executable, "-c", _PROBE], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        return {"error": (r.stderr or "").strip().splitlines()[-1:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_fire_estimate(cdft, info, threads, steps_gpu, steps=3):
    """Config 2 on the CPU: seconds per FIRE2 step of the oracle on the same slit pore (White Bear I + LJ QF
    mean field, walls, mu of rho_b = 0.7), times the GPU's step count to 1e-8: a labelled estimate."""
    import oracle
    oracle.set_threads(threads, threads)
    shape = tuple(info["grid"])
    n, dx, kT = shape[0], info["dx"], info["kT"]
    L = n * dx
    z = np.arange(n) * dx
    vz = 10.0 * (np.exp(-z / 0.2) + np.exp(-(L - z) / 0.2))
    vext = np.ascontiguousarray(np.broadcast_to(vz, shape))
    t0 = time.perf_counter()
    o = oracle.DFT(shape, dx, info["hsd"], mf=True, kT=kT, mu=info["mu"], vext=vext, accurate_weights=True)
    f = oracle.Fire(o, info["rho_bulk"] * np.exp(-vext))
    setup = time.perf_counter() - t0
    ts = []
    for _ in range(steps):
        t1 = time.perf_counter()
        f.step()
        ts.append(time.perf_counter() - t1)
    per = float(np.median(ts))
    return {"seconds_per_step": per, "steps_timed": steps, "setup_seconds": setup, "gpu_steps_to_tol": steps_gpu,
            "estimate_seconds_to_tol": per * steps_gpu, "threads": threads,
            "label": "estimate: CPU oracle seconds per FIRE2 step x the GPU run's step count to 1e-8"}


def cpu_legs(cdft, rho, threads, fire_info):
    """CPU baselines of the FP64 oracle: 256^3 with single-threaded FFTs (reference-faithful: the reference's FFTW
    plans run on one thread, include/DFT_Factory.h:86-90) and with threaded FFTs, the 32^3 config-1 case, and the
    config-2 FIRE estimate."""
    out = {"host": host_info(), "threads": threads,
           "fft_backend": "repo-written mixed-radix CPU FFT" + (" (fftw3.h present but unused)"
                                                                 if os.path.exists("/usr/include/fftw3.h") else
                                                                 " (FFTW3 is not installed)")}
    t_ii, ts_ii, F, dF = cpu_fmt(rho, SHAPE, DX, threads, threads, 3)
    out["fmt_256_threaded_fft"] = {"seconds": t_ii, "runs": ts_ii, "evals_per_s": 1.0 / t_ii,
                                   "pointwise_threads": threads, "fft_threads": threads}
    t_i, ts_i, _, _ = cpu_fmt(rho, SHAPE, DX, threads, 1, 1)
    out["fmt_256_reference_faithful"] = {"seconds": t_i, "runs": ts_i, "evals_per_s": 1.0 / t_i,
                                         "pointwise_threads": threads, "fft_threads": 1}
    shape1 = (32, 32, 32)
    u = splitmix64_uniform(32 ** 3, 0x5EED0001).reshape(shape1)
    rho1 = 0.6 * (1.0 + 0.05 * (2.0 * u - 1.0))
    for key, ft in (("fmt_32_reference_faithful", 1), ("fmt_32_threaded_fft", threads)):
        t, ts, _, _ = cpu_fmt(rho1, shape1, 0.1, threads, ft, 5)
        out[key] = {"seconds": t, "evals_per_s": 1.0 / t, "pointwise_threads": threads, "fft_threads": ft}
    if fire_info:
        out["config2_fire"] = cpu_fire_estimate(cdft, fire_info, threads, fire_info["steps"])
    return out, F, dF


def sources_sha16():
    """Digest of the HIP sources the library is built from (ties a measured profile to the kernels it measured)."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "classicaldft_amd", "csrc")
    for f in sorted(os.listdir(d)):
        if f.endswith((".hip", ".hpp", ".cpp")):
            with open(os.path.join(d, f), "rb") as fh:
                h.update(f.encode() + b"\0" + fh.read())
    return h.hexdigest()[:16]


def warm_start_leg(cdft, count, check_unet=True):
    """Learned warm start vs uniform start on held-out 64^3 Gaussian-well systems (FP64 FIRE to 1e-8), with the
    committed network; plus the network's own kernel throughput at 64^3 (forward batch 16, train step batch 8)."""
    from classicaldft_amd import warmstart as W
    wl = W.Workload(64)
    ws = W.WarmStart.load()
    out = W.compare(wl, ws, 0x5EED5000 + 4096, count, schedule="fp64")
    if check_unet:
        D = 64
        rng = np.random.default_rng(0)
        V = cdft.DeviceArray.from_host(rng.standard_normal((16, D, D, D)).astype(np.float32), np.float32)
        M = cdft.DeviceArray.from_host(np.zeros(16, np.float32), np.float32)
        T = cdft.DeviceArray.from_host(0.1 * rng.standard_normal((16, D, D, D)).astype(np.float32), np.float32)
        o = cdft.DeviceArray((16, D, D, D), np.float32)
        net = cdft.UNet(D, seed=3)
        idx16, idx8 = np.arange(16), np.arange(8)
        net.forward(V, M, idx16, o)
        t = time.perf_counter()
        for _ in range(10):
            net.forward(V, M, idx16, o)
        tf = (time.perf_counter() - t) / 10
        net.train_step(V, M, T, idx8, 1e-5)
        t = time.perf_counter()
        for _ in range(10):
            net.train_step(V, M, T, idx8, 1e-5)
        tt = (time.perf_counter() - t) / 10
        fwd = W.unet_flops(D)
        out["unet"] = {"grid": D, "forward_ms_batch16": tf * 1e3, "train_step_ms_batch8": tt * 1e3,
                       "forward_tflops": 16 * fwd / tf * 1e-12, "train_tflops": 8 * 3 * fwd / tt * 1e-12,
                       "flops_forward_per_sample": fwd, "peak_bf16_dense_tflops": 2500.0}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--fire-grid", type=int, default=128, help="slit-pore FIRE time-to-tolerance grid (0: skip)")
    ap.add_argument("--fire-max-steps", type=int, default=20000)
    ap.add_argument("--ddft-steps", type=int, default=0, help="DDFT fcc-seed steps at 256^3 (0: skip)")
    ap.add_argument("--batch", action="store_true", help="batched FIRE on 96 systems of 48^3 / 64^3 / 96^3")
    ap.add_argument("--no-ao", action="store_true", help="skip the Asakura-Oosawa polymer timing")
    ap.add_argument("--no-mix", action="store_true", help="skip the binary-mixture timing")
    ap.add_argument("--mixture-only", type=int, default=0,
                    help="time only this many binary-mixture evaluations and print their JSON (for profiling)")
    ap.add_argument("--warm-systems", type=int, default=64,
                    help="held-out 64^3 systems for the learned warm start vs uniform start (0: skip)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 8 * n, "read": 8 * n, "written": 8 * n}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the FMT evaluation of one grid "
              "is not sharded across GPUs here", file=sys.stderr)
        sys.exit(2)

    import classicaldft_amd as cdft

    rho = workload_density()
    rho_d = cdft.DeviceArray.from_host(rho)
    dF_d = cdft.DeviceArray(SHAPE)
    if args.mixture_only:
        print(json.dumps({"mixture_fp64": time_mixture(cdft, rho_d, args.mixture_only)}), flush=True)
        return
    print("# fp64 timing", flush=True)
    t64, F64, dF64, st64, lay64 = time_system(cdft, cdft.FP64, rho_d, dF_d, args.steps, args.warmup)
    print("# mixed timing", flush=True)
    tmx, Fmx, dFmx, stmx, laymx = time_system(cdft, cdft.MIXED, rho_d, dF_d, args.steps, args.warmup)

    ao = None
    if not args.no_ao:
        print("# AO polymer timing", flush=True)
        ao = time_ao(cdft, rho_d, dF_d, max(3, args.steps // 5))
        ao["vs_fp64_plain"] = ao["ms_per_step"] / t64

    mixture = None
    if not args.no_mix:
        print("# binary mixture timing", flush=True)
        mixture = time_mixture(cdft, rho_d, max(3, args.steps // 5))
        mixture["vs_fp64_one_species"] = mixture["ms_per_step"] / t64

    m = 11
    B64 = model_bytes(8, m)
    Bmx = model_bytes(4, m, fp64_io=True)
    ach = B64 / (t64 * 1e-3) / 1e9
    sb = stage_bytes(SHAPE, 8, lay64[0], lay64[1])
    dom = max(st64, key=st64.get) if st64 else None
    dominant = None
    if dom:
        dominant = {"kernel": dom, "ms": st64[dom], "bytes": sb.get(dom), "stored_arrays": lay64[0],
                    "achieved_GBs": (sb[dom] / (st64[dom] * 1e-3) / 1e9) if dom in sb else None,
                    "bytes_note": "compulsory bytes of this kernel: one pass over the arrays it reads and writes"}
    # measured HBM bytes per evaluation (rocprofv3 FETCH_SIZE / WRITE_SIZE passes, calibrated): used only when the
    # profile was measured on the kernels built from the current sources
    traffic, traffic_note = None, "no traffic profile"
    tpath = os.path.join(ROOT, "profiles", "traffic_fmt_256.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        if tj.get("sources_sha16") == sources_sha16():
            traffic = tj.get("hbm_bytes_per_eval_fp64")
            traffic_note = f"profiles/traffic_fmt_256.json, measured on these sources ({tj['sources_sha16']})"
        else:
            traffic_note = (f"profiles/traffic_fmt_256.json was measured on other kernel sources "
                            f"({tj.get('sources_sha16')}, current {sources_sha16()}): not used")

    print("# streaming probe", flush=True)
    probe = streaming_probe()
    result = {
        "metric": "FMT F+dF evals/s @256^3 (FP64 vs mixed); FIRE time-to-tol, warm vs cold",
        "value": 1000.0 / t64,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "fmt_f_df_whitebear_i_256^3", "grid": list(SHAPE), "dx": DX, "hsd": HSD,
                   "model": "WhiteBearI", "density": "0.6*(1+0.05*U[-1,1]) splitmix64 seed 0x5EED0256",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
                     "traffic": traffic, "traffic_source": traffic_note,
                     "scope": "one F+dF evaluation (all kernels), algorithmic bytes B(8,11)",
                     "algorithmic_bytes": B64, "dominant_kernel": dominant,
                     "streaming_probe": probe,
                     "frac_of_copy_rate": (ach / probe["copy_GBs"]) if "copy_GBs" in probe else None,
                     "traffic_rate_GBs": (traffic / (t64 * 1e-3) / 1e9) if traffic else None},
        "mixed": {"evals_per_s": 1000.0 / tmx, "ms_per_step": tmx, "speedup_vs_fp64": t64 / tmx,
                  "F_rel_err_vs_fp64": abs(Fmx - F64) / abs(F64),
                  "dF_maxrel_err_vs_fp64": float(np.abs(dFmx - dF64).max() / np.abs(dF64).max()),
                  "roofline_frac": Bmx / (tmx * 1e-3) / 1e9 / PEAK_HBM_GBS},
        "stages_ms_fp64": st64,
        "stages_ms_mixed": stmx,
        "ao_polymer_fp64": ao,
        "mixture_fp64": mixture,
        "F_fp64": F64,
    }
    if args.fire_grid:
        print("# FIRE time to tolerance", flush=True)
        result["fire_time_to_tol"] = {"slit_pore": fire_time_to_tol(cdft, n=args.fire_grid, max_steps=args.fire_max_steps),
                                      "slit_pore_mixed_then_fp64": fire_ladder_time_to_tol(
                                          cdft, n=args.fire_grid, max_steps=args.fire_max_steps),
                                      "slit_pore_device_control": fire_graph_time_to_tol(
                                          cdft, n=args.fire_grid, max_steps=args.fire_max_steps),
                                      "slit_pore_device_control_mixed_then_fp64": fire_graph_time_to_tol(
                                          cdft, n=args.fire_grid, ladder_tol=1e-5, max_steps=args.fire_max_steps),
                                      "start": "rho_bulk exp(-V_ext)", "tol": 1e-8,
                                      "criterion": "max_i |dF/drho_i| / dV (raw gradient)"}
    if args.warm_systems:
        print("# learned warm start vs uniform start", flush=True)
        result.setdefault("fire_time_to_tol", {})["warm_vs_cold"] = warm_start_leg(cdft, args.warm_systems)
    if args.batch:
        print("# batched FIRE", flush=True)
        result["batched_fire"] = batch_fire(cdft)
    if args.ddft_steps:
        print("# DDFT fcc seed", flush=True)
        result["ddft_fcc_seed"] = {"fp64": ddft_fcc_seed(cdft, args.ddft_steps, precision=cdft.FP64),
                                   "mixed": ddft_fcc_seed(cdft, args.ddft_steps, precision=cdft.MIXED)}
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(f"# cpu baseline ({threads} threads)", flush=True)
        slit = result.get("fire_time_to_tol", {}).get("slit_pore")
        legs, Fc, dFc = cpu_legs(cdft, rho, threads, slit)
        ref = legs["fmt_256_reference_faithful"]
        result["cpu_baseline"] = {
            "value": ref["evals_per_s"], "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": "one 256^3 White Bear I F+dF evaluation of the same density by the FP64 C++ oracle after a "
                      "warm-up, OpenMP pointwise loops on all the threads used and single-threaded FFTs as the "
                      "reference runs them; weight set-up excluded",
            "seconds": ref["seconds"], "gpu_speedup": (1000.0 / t64) / ref["evals_per_s"], **legs,
        }
        result["parity_256"] = {"F_rel": abs(F64 - Fc) / abs(Fc),
                                "dF_maxrel": float(np.abs(dF64 - dFc).max() / np.abs(dFc).max()),
                                "note": "GPU with device-generated weights vs oracle with accurate weights"}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
