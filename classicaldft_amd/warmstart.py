"""Learned warm start for FIRE (no counterpart in the reference).

A 3-level 3D U-Net (``cdft_unet_*``: BF16 MFMA convolutions with FP32
accumulation, trained on the device) maps (beta V_ext, beta mu) on a periodic
D^3 grid to u = log(rho / rho_b) + beta V_ext, rho_b the bulk density at mu
(so u vanishes where V_ext does and the level of the bulk is exact); the warm
start is rho0 = rho_b exp(u - beta V_ext).  FIRE (``FireBatch``) is the refiner and the judge: a member starts
from the network's density only when that density has a lower grand potential
than the uniform bulk density, and a member still running after
``fallback_steps`` FIRE steps is restarted from the uniform density.

The training set is made by the engine itself: FIRE equilibria of seeded
systems (``workloads.well_systems``), computed in batches on the device.

    python -m classicaldft_amd.warmstart --out weights.npz --report report.json

runs the whole pipeline (data, training, held-out evaluation against the
uniform start) and writes the FP32 parameters and a JSON report.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import cdft as C
from .workloads import well_systems

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warmstart_64.npz")


# (input channels, output channels, level) of the ten 3x3x3 convolutions, as in unet.hip
UNET_LAYERS = ((8, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2), (96, 32, 1),
               (32, 32, 1), (48, 16, 0), (16, 16, 0))


def unet_flops(D):
    """Multiply-add flops of one sample's forward convolutions (2 x 27 x cin x cout per output site; the
    backward data and weight products cost the same again each)."""
    return sum(2 * 27 * ci * co * (D >> lv) ** 3 for ci, co, lv in UNET_LAYERS)


def log(msg):
    print(msg, flush=True)


class Workload:
    """The physical set-up shared by data generation, training and evaluation: hard spheres (White Bear I)
    plus LJ mean field at kT = 1.5 epsilon, above the model's critical temperature (about 1.34), so that every
    bulk density in [0.3, 0.8] is a stable state and each system has one equilibrium.  (At kT = 0.8 the
    spinodal spans rho = 0.08 - 0.52: uniform starts phase-separate and the equilibrium is not a function of
    (V_ext, mu).)"""

    def __init__(self, D=64, dx=0.1, kT=1.5, model=C.WHITEBEAR_I):
        self.D, self.dx, self.kT = D, dx, kT
        self.kw = dict(hsd=C.lj_hsd(kT), model=model, ideal=True, mean_field=True, kT=kT)
        self._probe = C.System((D, D, D), dx, **self.kw)

    def systems(self, first_seed, count):
        """(beta mu, bulk densities, beta V_ext (float64)) of systems first_seed ..."""
        rho_b, V = well_systems(self.D, count, first_seed, self.dx)
        mus = np.array([self._probe.bulk(r)[1] for r in rho_b])
        return mus, rho_b, V

    def fire(self, mus, rho0, V, precision=C.FP64, **params):
        return C.FireBatch((self.D,) * 3, self.dx, mus, rho0, vexts=V, precision=precision, **self.kw, **params)


def run_to_tol(fb, tol, max_steps, check=25):
    """Run a FireBatch until every member has max |dF/drho| <= tol or max_steps; (reports, status, seconds)."""
    t = time.perf_counter()
    steps = 0
    while steps < max_steps:
        fb.run(check, tol)
        steps += check
        reps, st, running = fb.status()
        if running == 0:
            break
    return reps, st, time.perf_counter() - t


# FP64 throughout, or the precision ladder: mixed precision to 1e-5 (at most 600 steps), then FP64 from a FIRE
# restart (cdft_fire_batch_set_precision)
SCHEDULES = {"fp64": ((C.FP64, None, None),), "ladder": ((C.MIXED, 1e-5, 600), (C.FP64, None, None))}


def solve(fb, schedule, tol, max_steps):
    """Run the phases of a schedule (precision, phase tolerance or None for tol, step cap or None)."""
    t = time.perf_counter()
    for k, (prec, ptol, cap) in enumerate(SCHEDULES[schedule]):
        if k > 0:
            fb.set_precision(prec)
        reps, st, _ = run_to_tol(fb, tol if ptol is None else ptol, max_steps if cap is None else cap)
    return reps, st, time.perf_counter() - t


def equilibria(wl, first_seed, count, batch=256, tol=2e-5, max_steps=1500, precision=C.MIXED):
    """FIRE equilibria from the uniform bulk density, in device batches.
    Returns float32 arrays (beta V_ext, beta mu, u = log(rho / rho_b) + beta V_ext) and the converged mask."""
    D = wl.D
    Vs = np.empty((count, D, D, D), np.float32)
    U = np.empty((count, D, D, D), np.float32)
    M = np.empty(count, np.float32)
    ok = np.zeros(count, bool)
    for b0 in range(0, count, batch):
        nb = min(batch, count - b0)
        mus, rho_b, V = wl.systems(first_seed + b0, nb)
        rho0 = np.broadcast_to(rho_b[:, None, None, None], V.shape)
        fb = wl.fire(mus, rho0, V, precision=precision)
        fb.autotune(reps=2)
        reps, st, sec = run_to_tol(fb, tol, max_steps)
        for i in range(nb):
            rho = fb.density(i)
            U[b0 + i] = np.log(np.maximum(rho, 1e-30) / rho_b[i]) + V[i]
        Vs[b0:b0 + nb] = V
        M[b0:b0 + nb] = mus
        ok[b0:b0 + nb] = st == 1
        fb.close()
        log(f"# equilibria {b0 + nb}/{count}: batch {sec:.1f} s, converged {(st == 1).sum()}/{nb}, "
            f"mean steps {np.mean([r['steps'] for r in reps]):.0f}")
    return Vs, M, U, ok


def bulk_density(system, mu, hsd, lo=1e-8):
    """The bulk density whose beta mu (System.bulk) is `mu`, by bisection below packing fraction 0.74 (beta mu
    increases with density above the critical temperature)."""
    hi = 0.74 * 6.0 / (np.pi * hsd ** 3)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if system.bulk(mid)[1] < mu:
            lo = mid
        else:
            hi = mid
        if hi - lo <= 1e-15 * hi:
            break
    return 0.5 * (lo + hi)


class WarmStart:
    """The trained network and its use as a FIRE initial density."""

    def __init__(self, D=64, params=None, seed=1):
        self.net = C.UNet(D, seed)
        self.D = D
        if params is not None:
            self.net.set_params(params)

    @classmethod
    def load(cls, path=WEIGHTS):
        with np.load(path, allow_pickle=False) as z:
            p = z["params"].astype(np.float32)
            D = int(z["D"])
        return cls(D, p)

    def save(self, path, **meta):
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        np.savez_compressed(path, params=self.net.params(), D=self.D, **meta)

    def predict_u(self, V, mus, batch=16):
        """Network output u (float32) for host arrays V (n, D, D, D) and beta mu (n,)."""
        V = np.ascontiguousarray(V, np.float32)
        n = V.shape[0]
        dV = C.DeviceArray.from_host(V, np.float32)
        dmu = C.DeviceArray.from_host(np.asarray(mus, np.float32), np.float32)
        out = np.empty_like(V)
        bufs = {}  # one output buffer per batch size, reused
        for b0 in range(0, n, batch):
            nb = min(batch, n - b0)
            if nb not in bufs:
                bufs[nb] = C.DeviceArray((nb,) + V.shape[1:], np.float32)
            self.net.forward(dV, dmu, np.arange(b0, b0 + nb), bufs[nb])
            out[b0:b0 + nb] = bufs[nb].download()
        for o in bufs.values():
            o.free()
        dV.free()
        dmu.free()
        return out

    def predict(self, V, mus, rho_b):
        """Warm-start densities rho0 = rho_b exp(u - beta V_ext) (float64); rho_b: bulk densities at mu
        (``bulk_density``).  Formed in place in float32 (a starting point for FIRE; its rounding is far below the
        network's error) and widened once."""
        u = self.predict_u(V, mus)
        np.subtract(u, V, out=u, casting="unsafe")
        np.exp(u, out=u)
        u *= np.asarray(rho_b, np.float32)[:, None, None, None]
        np.minimum(u, np.float32(7.0), out=u)  # far above any fluid density
        return u.astype(np.float64)


def train(net, V, mus, U, train_idx, val_idx, epochs=24, batch=8, lr=2e-3, lr_min=2e-5, seed=7, budget_s=None):
    """Adam with a cosine learning-rate schedule on the device-resident data set; returns the history."""
    dV = C.DeviceArray.from_host(V, np.float32)
    dM = C.DeviceArray.from_host(mus, np.float32)
    dU = C.DeviceArray.from_host(U, np.float32)
    # start the head at the mean target so the first steps fit structure, not an offset
    p = net.params()
    p[-1] = float(np.mean(U[train_idx[:64]]))
    net.set_params(p)
    rng = np.random.default_rng(seed)
    per_epoch = len(train_idx) // batch
    total = epochs * per_epoch
    hist = []
    t0 = time.perf_counter()
    k = 0
    for ep in range(epochs):
        order = rng.permutation(train_idx)
        run = 0.0
        for j in range(per_epoch):
            a = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * k / total))
            run += net.train_step(dV, dM, dU, order[j * batch:(j + 1) * batch], a)
            k += 1
        vl = validation_loss(net, dV, dM, U, val_idx)
        hist.append({"epoch": ep + 1, "train_loss": run / per_epoch, "val_loss": vl,
                     "seconds": time.perf_counter() - t0})
        log(f"# epoch {ep + 1}: train {run / per_epoch:.3e} val {vl:.3e} ({time.perf_counter() - t0:.0f} s)")
        if budget_s is not None and time.perf_counter() - t0 > budget_s:
            break
    for a in (dV, dM, dU):
        a.free()
    return hist


def validation_loss(net, dV, dM, U, idx, batch=16):
    tot = 0.0
    for b0 in range(0, len(idx), batch):
        ii = idx[b0:b0 + batch]
        o = C.DeviceArray((len(ii),) + U.shape[1:], np.float32)
        net.forward(dV, dM, ii, o)
        tot += float(np.sum((o.download().astype(np.float64) - U[ii]) ** 2))
        o.free()
    return tot / (len(idx) * np.prod(U.shape[1:]))


def grand_potentials(wl, mus, V, rho):
    """beta Omega of each member (inf where a weighted density leaves the physical range)."""
    sysb = C.System((wl.D,) * 3, wl.dx, mus=mus, vexts=V, **wl.kw)
    try:
        parts, _ = sysb.eval(rho)
        om = parts[:, 0].copy()
    except C.CdftError:
        om = np.full(len(mus), np.inf)
        for i in range(len(mus)):
            s1 = C.System((wl.D,) * 3, wl.dx, mu=mus[i], vext=V[i], **wl.kw)
            try:
                om[i] = s1.eval(rho[i])[0][0]
            except C.CdftError:
                pass
            s1.close()
    sysb.close()
    om[~np.isfinite(om)] = np.inf
    return om


def uniform_grand_potentials(wl, mus, rho_b, V):
    """beta Omega of the uniform bulk density in each member's field, in closed form: at uniform density the
    lattice functional is the bulk one (the weights' sum rules give the bulk measures, the kernel sums to a_vdw),
    so Omega = V_box (f(rho_b) - mu rho_b) + rho_b dV sum_x V(x)."""
    dV = wl.dx ** 3
    f = np.array([wl._probe.bulk(r)[0] for r in rho_b])
    vsum = V.reshape(len(mus), -1).sum(axis=1, dtype=np.float64)
    return wl.D ** 3 * dV * (f - mus * rho_b) + rho_b * dV * vsum


def judged_batch(wl, ws, mus, rho_b, V, precision=C.FP64):
    """A FireBatch started from the network's densities where FIRE's objective (the grand potential, from the
    batch's own first evaluation) is lower there than at the uniform bulk density (closed form), from the uniform
    density elsewhere.  Returns (batch, used_warm mask, seconds spent on the network and the judgement)."""
    t = time.perf_counter()
    warm = ws.predict(V, mus, rho_b)
    fb = wl.fire(mus, warm, V, precision=precision)
    reps, st, _ = fb.status()
    om_w = np.array([r["F"] if s == 0 else np.inf for r, s in zip(reps, st)])
    use = om_w < uniform_grand_potentials(wl, mus, rho_b, V)
    tj = time.perf_counter() - t
    if not use.all():
        fb.close()
        uni = np.broadcast_to(rho_b[:, None, None, None], V.shape)
        fb = wl.fire(mus, np.where(use[:, None, None, None], warm, uni), V, precision=precision)
    return fb, use, tj


def compare(wl, ws, first_seed, count, tol=1e-8, max_steps=20000, fallback_steps=None, schedule="fp64"):
    """FIRE to max |dF/drho| <= tol from the uniform bulk density and from the judged warm start on held-out
    systems.  Both wall times include creating the FIRE batch (its first evaluation); the warm one also the
    network and the judgement.  Members of the warm run still running after fallback_steps (default: the most
    steps any uniform start needed) restart from the uniform density in a second batch."""
    mus, rho_b, V = wl.systems(first_seed, count)
    uni = np.ascontiguousarray(np.broadcast_to(rho_b[:, None, None, None], V.shape))
    # untimed warm-up of the network and of both precisions at this size
    ws.predict(V[:2], mus[:2], rho_b[:2])
    fb = wl.fire(mus, uni, V, precision=SCHEDULES[schedule][0][0])
    solve(fb, schedule, 1.0, 25)
    fb.close()

    t0 = time.perf_counter()
    fb = wl.fire(mus, uni, V, precision=SCHEDULES[schedule][0][0])
    rc, sc, _ = solve(fb, schedule, tol, max_steps)
    tc = time.perf_counter() - t0
    fb.close()
    steps_c = np.array([r["steps"] for r in rc])
    if fallback_steps is None:
        fallback_steps = int(np.max(steps_c))

    t0 = time.perf_counter()
    fb, used, tj = judged_batch(wl, ws, mus, rho_b, V, precision=SCHEDULES[schedule][0][0])
    rw, sw, _ = solve(fb, schedule, tol, fallback_steps)
    steps_w = np.array([r["steps"] for r in rw])
    F_w = np.array([r["F"] for r in rw])
    fb.close()
    late = np.where(sw != 1)[0]
    if late.size:
        fb = wl.fire(mus[late], uni[late], V[late], precision=SCHEDULES[schedule][0][0])
        rl, sl, _ = solve(fb, schedule, tol, max_steps)
        fb.close()
        for j, i in enumerate(late):
            steps_w[i] += rl[j]["steps"]
            F_w[i] = rl[j]["F"]
            sw[i] = sl[j]
    tw = time.perf_counter() - t0
    F_c = np.array([r["F"] for r in rc])
    return {
        "systems": count, "first_seed": first_seed, "tol": tol, "schedule": schedule,
        "cold": {"seconds": tc, "mean_steps": float(steps_c.mean()), "max_steps": int(steps_c.max()),
                 "converged": int((sc == 1).sum())},
        "warm": {"seconds": tw, "judge_seconds": tj, "mean_steps": float(steps_w.mean()),
                 "max_steps": int(steps_w.max()), "converged": int((sw == 1).sum()),
                 "judged_uniform": int((~used).sum()), "fallback_after_steps": fallback_steps,
                 "restarted_uniform": int(late.size)},
        "wall_speedup": tc / tw, "step_speedup": float(steps_c.mean() / steps_w.mean()),
        "fallback_rate": float(((~used).sum() + late.size) / count),
        "same_minimum": int(np.sum(np.abs(F_w - F_c) <= 1e-8 * np.abs(F_c))),
        "timing": "both wall times include creating the FIRE batch; warm also the network and the judgement",
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--D", type=int, default=64)
    ap.add_argument("--kT", type=float, default=1.5)
    ap.add_argument("--train", type=int, default=4096, help="training systems (seeds 0x5EED5000 + i)")
    ap.add_argument("--heldout", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=24)
    ap.add_argument("--budget", type=float, default=None, help="training time limit in seconds")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--gen-batch", type=int, default=256)
    ap.add_argument("--gen-tol", type=float, default=2e-5)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    seed0 = 0x5EED5000
    wl = Workload(a.D, kT=a.kT)
    t = time.perf_counter()
    V, M, U, ok = equilibria(wl, seed0, a.train, batch=a.gen_batch, tol=a.gen_tol)
    t_gen = time.perf_counter() - t
    idx = np.where(ok)[0]
    val = idx[-64:]
    tr = idx[:-64]
    log(f"# data: {len(idx)} converged of {a.train} in {t_gen:.0f} s; training on {len(tr)}")
    ws = WarmStart(a.D, seed=1)
    hist = train(ws.net, V, M, U, tr, val, epochs=a.epochs, batch=a.batch, lr=a.lr, budget_s=a.budget)
    ws.save(a.out, train_systems=len(tr), first_seed=seed0, kT=a.kT, dx=wl.dx, target="log(rho/rho_b)+beta*V")
    del V, U
    rep = {"data_seconds": t_gen, "train_systems": int(len(tr)), "history": hist}
    for sched in SCHEDULES:
        r = compare(wl, ws, seed0 + a.train, a.heldout, schedule=sched)
        log(json.dumps(r))
        rep["compare_" + sched] = r
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
