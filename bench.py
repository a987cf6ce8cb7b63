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
m = 11; the dominant kernel's own compulsory bytes are reported beside it.

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

SEED = 0x5EED0256
SHAPE = (256, 256, 256)
DX = 0.04
HSD = 1.0
PEAK_HBM_GBS = 8000.0


def splitmix64_uniform(n, seed):
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


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


def stage_bytes(shape, e, m):
    """Compulsory HBM bytes of each kernel of this implementation (one pass each)."""
    Nx, Ny, Nz = shape
    N = Nx * Ny * Nz
    Nk = Nx * Ny * (Nz // 2 + 1)
    c = 2 * e
    return {
        "z_r2c": 8 * N + c * Nk,
        "y_fwd": 2 * c * Nk,
        "x_conv_fwd": c * Nk + m * e * Nk + m * c * Nk,
        "y_inv": 2 * m * c * Nk,
        "z_pointwise": 2 * m * c * Nk,
        "y_fwd2": 2 * m * c * Nk,
        "x_conv_bwd": m * c * Nk + m * e * Nk + c * Nk,
        "y_inv2": 2 * c * Nk,
        "z_final": c * Nk + 8 * N + 8 * N,  # acc in, rho in, dF out
    }


def time_system(cdft, precision, rho_d, dF_d, steps, warmup):
    s = cdft.System(SHAPE, DX, HSD, model=cdft.WHITEBEAR_I, precision=precision)
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
    return ms / steps, parts[2], dF, acc


def cpu_baseline(rho, threads):
    import oracle
    oracle.set_threads(threads, threads)
    t0 = time.time()
    h = oracle.FMT(oracle.WBI, SHAPE, DX, HSD, weights=oracle.fmt_weights(SHAPE, DX, HSD,
                                                                          mode=oracle.WEIGHTS_ACCURATE))
    setup = time.time() - t0
    t0 = time.time()
    F, dF = h.eval(rho)
    dt = time.time() - t0
    return dt, setup, F, dF


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
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
    print("# fp64 timing", flush=True)
    t64, F64, dF64, st64 = time_system(cdft, cdft.FP64, rho_d, dF_d, args.steps, args.warmup)
    print("# mixed timing", flush=True)
    tmx, Fmx, dFmx, stmx = time_system(cdft, cdft.MIXED, rho_d, dF_d, args.steps, args.warmup)

    m = 11
    B64 = model_bytes(8, m)
    Bmx = model_bytes(4, m, fp64_io=True)
    ach = B64 / (t64 * 1e-3) / 1e9
    sb = stage_bytes(SHAPE, 8, m)
    dom = max(st64, key=st64.get) if st64 else None
    dominant = None
    if dom:
        dominant = {"kernel": dom, "ms": st64[dom], "bytes": sb.get(dom),
                    "achieved_GBs": (sb[dom] / (st64[dom] * 1e-3) / 1e9) if dom in sb else None}
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_fmt_256.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_eval_fp64")

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
                     "traffic": traffic, "scope": "one F+dF evaluation (all kernels), algorithmic bytes B(8,11)",
                     "algorithmic_bytes": B64, "dominant_kernel": dominant},
        "mixed": {"evals_per_s": 1000.0 / tmx, "ms_per_step": tmx, "speedup_vs_fp64": t64 / tmx,
                  "F_rel_err_vs_fp64": abs(Fmx - F64) / abs(F64),
                  "dF_maxrel_err_vs_fp64": float(np.abs(dFmx - dF64).max() / np.abs(dF64).max()),
                  "roofline_frac": Bmx / (tmx * 1e-3) / 1e9 / PEAK_HBM_GBS},
        "stages_ms_fp64": st64,
        "stages_ms_mixed": stmx,
        "F_fp64": F64,
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(f"# cpu baseline ({threads} threads)", flush=True)
        dt, setup, Fc, dFc = cpu_baseline(rho, threads)
        result["cpu_baseline"] = {
            "value": 1.0 / dt, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": "one 256^3 White Bear I F+dF evaluation of the same density by the FP64 C++ oracle "
                      "(OpenMP pointwise loops and FFT rows, repo-written CPU FFT; weight set-up excluded, "
                      f"{setup:.1f} s)",
            "seconds": dt,
            "gpu_speedup": (1.0 / dt and (1000.0 / t64) / (1.0 / dt)),
        }
        result["parity_256"] = {"F_rel": abs(F64 - Fc) / abs(Fc),
                                "dF_maxrel": float(np.abs(dF64 - dFc).max() / np.abs(dFc).max()),
                                "note": "GPU with device-generated weights vs oracle with accurate weights"}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
