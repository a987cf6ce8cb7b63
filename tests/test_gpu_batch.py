"""Batched systems (one launch per kernel for all members), device-side FIRE2
over a batch (direct and hipGraph-captured), and the tile autotuner.

Each member of a batch must give exactly what a single system with the same
inputs gives: the evaluation runs the same kernels with a member offset, and
the batched FIRE2 runs the reference step (src/Minimizer.cpp:160-327) with its
control logic on the device, checked step-for-step against the host-controlled
cdft_fire_step trajectory (itself replayed against the oracle in
test_gpu_solvers.py).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SHAPE, DX, KT = (16, 12, 24), 0.1, 1.0


def member_inputs(nb, shape=SHAPE, dx=DX, seed=0x5EED4100):
    """Per member: bulk density, a Gaussian-well external field, a perturbed start density."""
    N = int(np.prod(shape))
    L = np.array(shape) * dx
    grid = np.meshgrid(*[np.arange(n) * dx for n in shape], indexing="ij")
    rhos, vexts, rho_b = [], [], []
    for b in range(nb):
        u = oracle.splitmix64_uniform(16 + N, seed + b)
        rb = 0.3 + 0.5 * u[0]
        c = u[1:4] * L
        v = np.zeros(shape)
        for k in range(3):
            d2 = sum(np.minimum(np.abs(g - c[i]), L[i] - np.abs(g - c[i])) ** 2 for i, g in enumerate(grid))
            v += (6 * u[4 + k] - 3) * np.exp(-d2 / (2 * (0.3 + 0.7 * u[7 + k]) ** 2))
        rhos.append(rb * np.exp(-0.3 * v) * (1 + 0.05 * (2 * u[16:].reshape(shape) - 1)))
        vexts.append(v)
        rho_b.append(rb)
    return np.array(rho_b), np.array(vexts), np.array(rhos)


def test_batch_eval_matches_single_systems(cdft):
    nb = 3
    rho_b, vexts, rhos = member_inputs(nb)
    mus = -1.0 + 0.5 * np.arange(nb)
    kw = dict(hsd=1.0, model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=KT)
    bs = cdft.System(SHAPE, DX, mus=mus, vexts=vexts, **kw)
    parts, dF = bs.eval(rhos)
    for b in range(nb):
        s = cdft.System(SHAPE, DX, mu=mus[b], vext=vexts[b], **kw)
        p1, d1 = s.eval(rhos[b])
        assert np.max(np.abs(parts[b] - p1)) <= 1e-13 * np.max(np.abs(p1))
        assert np.max(np.abs(dF[b] - d1)) <= 1e-13 * np.max(np.abs(d1))
    F_hs, dfh = bs.fmt_eval(rhos)
    for b in range(nb):
        F1, d1 = cdft.System(SHAPE, DX, **kw).fmt_eval(rhos[b])
        assert abs(F_hs[b] - F1) <= 1e-13 * abs(F1)


@pytest.mark.parametrize("graph", [False, True])
def test_fire_batch_matches_host_fire(cdft, graph):
    nb, steps = 3, 30
    rho_b, vexts, rhos = member_inputs(nb, seed=0x5EED4200)
    mus = np.array([cdft.System(SHAPE, DX, hsd=1.0, ideal=True, mean_field=True, kT=KT).bulk(r)[1] for r in rho_b])
    kw = dict(hsd=1.0, model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=KT)
    fb = cdft.FireBatch(SHAPE, DX, mus, rhos, vexts=vexts, **kw)
    for _ in range(3):
        fb.run(steps // 3, 0.0, graph=graph)
    reps, status, running = fb.status()
    assert running == nb and (status == 0).all()
    for b in range(nb):
        s = cdft.System(SHAPE, DX, mu=mus[b], vext=vexts[b], **kw)
        f = cdft.Fire(s, rhos[b])
        for _ in range(steps):
            rc, info = f.step()
        rho_h, *_ = f.state()
        r = reps[b]
        assert r["steps"] == info["steps"] == steps
        assert r["calls"] == info["calls"]
        assert abs(r["F"] - info["F"]) <= 1e-10 * abs(info["F"])
        assert abs(r["fmax_raw"] - info["fmax_raw"]) <= 1e-8 * info["fmax_raw"]
        assert abs(r["dt"] - info["dt"]) <= 1e-12 * info["dt"]
        assert np.max(np.abs(fb.density(b) - rho_h)) <= 1e-10 * np.max(rho_h)
        # the whole FIRE state of the member (cdft_fire_batch_state) against the host-controlled run's
        rho_s, x_s, v_s, df_s = fb.state(b)
        rho_h, x_h, v_h, df_h = f.state()
        assert np.array_equal(rho_s, fb.density(b))
        for got, want in ((df_s, df_h), (x_s, x_h), (v_s, v_h)):
            assert np.max(np.abs(got - want)) <= 1e-9 * np.max(np.abs(want)) + 1e-300


def test_fire_batch_backtracks_like_host_fire(cdft):
    """A huge time step drives eta > 1: the device control restores x, halves v and dt exactly as the host
    control does (Minimizer.cpp:233-245)."""
    nb, shape = 2, (16, 16, 16)
    rhos = np.array([0.8 * (1 + 0.3 * (2 * oracle.splitmix64_uniform(4096, 5 + b).reshape(shape) - 1))
                     for b in range(nb)])
    mus = np.array([3.0, 2.5])
    kw = dict(hsd=1.0, model=cdft.WHITEBEAR_I, ideal=True)
    fb = cdft.FireBatch(shape, 0.1, mus, rhos, dt=5.0, dt_max=50.0, **kw)
    fb.run(6, 0.0, graph=True)
    reps, status, _ = fb.status()
    seen = 0
    for b in range(nb):
        f = cdft.Fire(cdft.System(shape, 0.1, mu=mus[b], **kw), rhos[b], dt=5.0, dt_max=50.0)
        for _ in range(6):
            rc, info = f.step()
            seen += rc == 1
        assert reps[b]["calls"] == info["calls"]
        assert reps[b]["steps"] == info["steps"]
        assert abs(reps[b]["dt"] - info["dt"]) <= 1e-12 * info["dt"]
        assert abs(reps[b]["F"] - info["F"]) <= 1e-10 * abs(info["F"])
    assert seen >= 1, "the test never drove eta above 1"


def test_fire_batch_converges_and_stops(cdft):
    nb = 2
    rho_b, vexts, rhos = member_inputs(nb, seed=0x5EED4400)
    kw = dict(hsd=1.0, model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=KT)
    mus = np.array([cdft.System(SHAPE, DX, hsd=1.0, ideal=True, mean_field=True, kT=KT).bulk(r)[1] for r in rho_b])
    fb = cdft.FireBatch(SHAPE, DX, mus, rhos, vexts=vexts, **kw)
    tol = 1e-8
    for _ in range(200):
        fb.run(25, tol)
        reps, status, running = fb.status()
        if running == 0:
            break
    assert (status == 1).all(), status
    for b in range(nb):
        assert reps[b]["fmax_raw"] <= tol
    steps = [r["steps"] for r in reps]
    fb.run(25, tol)
    reps2, _, _ = fb.status()
    assert [r["steps"] for r in reps2] == steps  # converged members do not step


def test_autotune_keeps_results(cdft):
    nb = 2
    rho_b, vexts, rhos = member_inputs(nb, shape=(32, 24, 48), seed=0x5EED4500)
    kw = dict(hsd=1.0, model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=KT)
    bs = cdft.System((32, 24, 48), DX, mus=[0.1, 0.2], vexts=vexts, **kw)
    p0, d0 = bs.eval(rhos)
    ms = bs.autotune(reps=2)
    assert ms > 0
    p1, d1 = bs.eval(rhos)
    assert np.max(np.abs(p1 - p0)) <= 1e-13 * np.max(np.abs(p0))
    assert np.max(np.abs(d1 - d0)) <= 1e-13 * np.max(np.abs(d0))
    s = cdft.System((32, 24, 48), DX, mu=0.1, vext=vexts[0], **kw)
    q0, e0 = s.eval(rhos[0])
    assert s.autotune(reps=2) > 0
    q1, e1 = s.eval(rhos[0])
    assert np.max(np.abs(q1 - q0)) <= 1e-13 * np.max(np.abs(q0))
    assert np.max(np.abs(e1 - e0)) <= 1e-13 * np.max(np.abs(e0))


def test_precision_ladder_reaches_fp64_equilibrium(cdft):
    """Mixed precision until its gradient noise floor, then FP64: the same equilibrium as FP64 throughout."""
    nb = 2
    rho_b, vexts, rhos = member_inputs(nb, seed=0x5EED4600)
    kw = dict(hsd=1.0, model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=KT)
    mus = np.array([cdft.System(SHAPE, DX, hsd=1.0, ideal=True, mean_field=True, kT=KT).bulk(r)[1] for r in rho_b])
    ref = cdft.FireBatch(SHAPE, DX, mus, rhos, vexts=vexts, **kw)
    lad = cdft.FireBatch(SHAPE, DX, mus, rhos, vexts=vexts, precision=cdft.MIXED, **kw)
    for fb, phases in ((ref, [(cdft.FP64, 1e-9)]), (lad, [(cdft.MIXED, 3e-6), (cdft.FP64, 1e-9)])):
        for prec, tol in phases:
            fb.set_precision(prec)
            for _ in range(400):
                fb.run(25, tol)
                _, st, running = fb.status()
                if running == 0:
                    break
            assert (st == 1).all(), (prec, st)
    # both end at max |dF/drho| <= 1e-9 of the FP64 functional; the grand potentials agree to second order in
    # the density difference, which a soft mode of this near-critical LJ state leaves at ~1e-6
    r_ref, _, _ = ref.status()
    r_lad, _, _ = lad.status()
    for b in range(nb):
        assert abs(r_lad[b]["F"] - r_ref[b]["F"]) <= 1e-10 * abs(r_ref[b]["F"])
        r0, r1 = ref.density(b), lad.density(b)
        assert np.max(np.abs(r1 - r0)) <= 1e-5 * np.max(r0)
