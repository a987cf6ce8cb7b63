"""GPU FIRE2 and DDFT against the oracle's step-by-step restatement.

Trajectory replay: the same initial density, the same functional and the
reference control logic on both sides; per-step F, time step, alpha and the
force norms agree to 1e-10 relative, and the final densities to 1e-10.
Invariants: DDFT conserves mass to 1e-12 (k = 0 mode), and a small-amplitude
ideal-gas mode decays at the discrete diffusion rate.  A forced eta > 1 step
back-tracks identically on both sides.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def density(shape, rho_b, seed, amp):
    u = oracle.splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    return rho_b * (1.0 + amp * (2.0 * u - 1.0))


def close(a, b, tol):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300)


def test_bulk_thermodynamics(cdft):
    shape, dx = (32, 32, 32), 0.1
    s = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=0.8)
    o = oracle.DFT(shape, dx, 1.0, mf=True, kT=0.8)
    for rho in (0.1, 0.5, 0.7, 0.9):
        f, mu = s.bulk(rho)
        fex, muex = oracle.bulk(oracle.WBI, rho, 1.0)
        # a_vdw itself is a sum over ~10^4 kernel cells in a different order on
        # each side (agreement ~1e-13), hence 1e-12 here
        assert close(s.a_vdw, o.a_vdw, 1e-12)
        assert close(mu, np.log(rho) + muex + o.a_vdw * rho, 1e-12)
        assert close(f, rho * np.log(rho) - rho + fex + 0.5 * o.a_vdw * rho ** 2, 1e-12)
        # thermodynamic consistency: mu = df/drho
        h = 1e-6 * rho
        fp, _ = s.bulk(rho + h)
        fm, _ = s.bulk(rho - h)
        assert close((fp - fm) / (2 * h), mu, 1e-8)


@pytest.mark.parametrize("mf", [False, True])
def test_fire_trajectory_replay(cdft, mf):
    shape, dx, hsd = (16, 16, 16), 0.1, 1.0
    rho0 = density(shape, 0.5, 11, 0.1)
    w = oracle.fmt_weights(shape, dx, hsd)
    kT = 1.0
    o = oracle.DFT(shape, dx, hsd, mf=mf, kT=kT, mu=0.5)
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mu=0.5, mean_field=mf, kT=kT)
    fo = oracle.Fire(o, rho0)
    fg = cdft.Fire(s, rho0)
    F0 = o.eval(rho0)[0][0]
    for i in range(40):
        rc0, a = fo.step()
        rc1, b = fg.step()
        assert rc0 == rc1, (i, rc0, rc1)
        assert close(a["F"], b["F"], 1e-10), (i, a["F"], b["F"])
        assert close(a["dt"], b["dt"], 1e-12) and close(a["alpha"], b["alpha"], 1e-12)
        assert close(a["fmax"], b["fmax_alias"], 1e-9) and close(a["fmax_raw"], b["fmax_raw"], 1e-9)
    r0 = fo.state()[0]
    r1 = fg.state()[0]
    assert np.abs(r0 - r1).max() <= 1e-10 * np.abs(r0).max()
    assert b["F"] < F0  # the free energy went down from the start


def test_fire_backtrack_on_eta_too_large(cdft):
    """A huge time step drives eta > 1: both sides restore x, halve v and dt."""
    shape, dx, hsd = (16, 16, 16), 0.1, 1.0
    rho0 = density(shape, 0.8, 5, 0.3)
    w = oracle.fmt_weights(shape, dx, hsd)
    o = oracle.DFT(shape, dx, hsd, mu=3.0)
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mu=3.0)
    fo = oracle.Fire(o, rho0, dt=5.0, dt_max=50.0)
    fg = cdft.Fire(s, rho0, dt=5.0, dt_max=50.0)
    seen = 0
    for i in range(6):
        rc0, a = fo.step()
        rc1, b = fg.step()
        assert rc0 == rc1
        assert close(a["dt"], b["dt"], 1e-12)
        seen += rc0 == 1
    assert seen >= 1, "the test never drove eta above 1"


def test_ddft_replay_and_mass(cdft):
    shape, dx, hsd = (16, 16, 16), 0.1, 1.0
    rho0 = density(shape, 0.5, 21, 0.1)
    w = oracle.fmt_weights(shape, dx, hsd)
    o = oracle.DFT(shape, dx, hsd, mf=True, kT=1.0)
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mean_field=True, kT=1.0)
    do = oracle.DDFT(o, rho0)
    dg = cdft.DDFT(s, rho0)
    m0 = rho0.sum() * dx ** 3
    for i in range(4):
        a = do.step()
        b = dg.step()
        assert a["calls"] == b["calls"] and a["restarts"] == b["restarts"]
        assert close(a["F"], b["F"], 1e-10) and close(a["dt"], b["dt"], 1e-12)
        assert abs(b["mass"] / m0 - 1) < 1e-12
    r0, r1 = do.density(), dg.density()
    assert np.abs(r0 - r1).max() <= 1e-10 * np.abs(r0).max()


def test_ddft_central_differences_replay(cdft):
    """DDFT(dft, showGraphics, central_differences = true) (include/Minimizer.h:192): the excess RHS through the
    central-difference g . dF (src/DDFT.cpp:320-351); the device replays the oracle step for step."""
    shape, dx, hsd = (16, 16, 16), 0.1, 1.0
    rho0 = density(shape, 0.5, 23, 0.1)
    w = oracle.fmt_weights(shape, dx, hsd)
    o = oracle.DFT(shape, dx, hsd, mf=True, kT=1.0)
    o.set_central_differences(True)
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mean_field=True, kT=1.0)
    do = oracle.DDFT(o, rho0)
    dg = cdft.DDFT(s, rho0, central_differences=True)
    dp = cdft.DDFT(s, rho0)
    m0 = rho0.sum() * dx ** 3
    for i in range(3):
        a = do.step()
        b = dg.step()
        dp.step()
        assert a["calls"] == b["calls"] and a["restarts"] == b["restarts"]
        assert close(a["F"], b["F"], 1e-10) and close(a["dt"], b["dt"], 1e-12)
        assert abs(b["mass"] / m0 - 1) < 1e-12
    r0, r1 = do.density(), dg.density()
    assert np.abs(r0 - r1).max() <= 1e-10 * np.abs(r0).max()
    assert np.abs(dp.density() - r1).max() > 1e-6 * np.abs(r0).max()  # the stride matters


def test_ddft_mean_field_linear_response(cdft):
    """The linear-response pin of tests/test_oracle_ddft.py on the device: a density mode of the ideal gas with a
    mean-field kernel decays at Lambda_k (1 + rho0 w_k), to second order in dt, and the device replays the oracle."""
    import test_oracle_ddft as lr
    w, rate, mode = lr.mode_setup()
    errs = []
    for dt in (4e-3, 2e-3):
        s = cdft.System(lr.SHAPE, lr.DX, 1.0, model=cdft.MODEL_NONE, ideal=True, mean_field=True, mf_kernel=w)
        d = cdft.DDFT(s, lr.RHO0 * (1 + lr.EPS * mode), dt=dt, dt_max=dt, tol=1e-14)
        for _ in range(int(round(lr.T / dt))):
            r = d.step()
        assert r["restarts"] == 0
        e = lr.amplitude(d.density(), mode) / (lr.EPS * np.exp(rate * r["time"])) - 1
        e0, _ = lr.oracle_run(dt, 1e-14)
        assert abs(e - e0) <= 1e-7  # rounding of rho (~1e-16) on a mode of amplitude 5e-7
        errs.append(e)
    assert abs(errs[0]) < 5e-5 and 3.6 < errs[0] / errs[1] < 4.4


def test_ddft_ideal_gas_mode_decay(cdft):
    """Ideal gas, one small Fourier mode: rho_k(t) = rho_k(0) exp(Lambda_k t) to O(amplitude)."""
    shape, dx = (16, 16, 16), 0.1
    x = np.arange(16)
    amp = 1e-7
    rho0 = 0.5 * (1 + amp * np.cos(2 * np.pi * 2 * x / 16))[:, None, None] * np.ones(shape)
    s = cdft.System(shape, dx, 1.0, model=cdft.MODEL_NONE, ideal=True)
    d = cdft.DDFT(s, rho0, dt=1e-4, dt_max=1e-4)
    for _ in range(5):
        r = d.step()
    rho = d.density()
    lam = 2 / dx ** 2 * (np.cos(2 * np.pi * 2 / 16) - 1)
    c_meas = (rho[:, 0, 0] - 0.5) @ np.cos(2 * np.pi * 2 * x / 16) / 8
    c_exp = 0.5 * amp * np.exp(lam * r["time"])
    assert abs(c_meas / c_exp - 1) < 1e-5


def test_fire_fixed_direction_and_reset_replay(cdft):
    """Minimizer::set_fixed_direction and fireMinimizer2::reset (src/Minimizer.cpp:30-34, 103-110, 149-158)."""
    shape, dx, hsd = (16, 16, 16), 0.1, 1.0
    rho0 = density(shape, 0.5, 13, 0.1)
    w = oracle.fmt_weights(shape, dx, hsd)
    e = np.cos(2 * np.pi * np.arange(16) / 16)[:, None, None] * np.ones(shape)
    e /= np.linalg.norm(e)
    o = oracle.DFT(shape, dx, hsd, mf=True, kT=1.0, mu=0.5)
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mu=0.5, mean_field=True, kT=1.0)
    fo, fg = oracle.Fire(o, rho0), cdft.Fire(s, rho0)
    for f in (fo, fg):
        f.set_fixed_direction(e)
        assert f.reset(rho0) == 0
    for i in range(24):
        if i == 12:  # restart both from the current density, dt and alpha kept
            for f in (fo, fg):
                f.set_fixed_direction(None)
                assert f.reset(fo.state()[0]) == 0
        rc0, a = fo.step()
        rc1, b = fg.step()
        assert rc0 == rc1 == 0
        assert close(a["F"], b["F"], 1e-10), (i, a["F"], b["F"])
        assert close(a["dt"], b["dt"], 1e-12) and close(a["alpha"], b["alpha"], 1e-12)
    r0, r1 = fo.state()[0], fg.state()[0]
    assert np.abs(r0 - r1).max() <= 1e-10 * np.abs(r0).max()
    _, _, _, d1 = fg.state()
    _, _, _, d0 = fo.state()
    assert np.abs(d0 - d1).max() <= 1e-9 * np.abs(d0).max()


def test_ddft_stalls_instead_of_looping(cdft):
    """A fixed point that never converges (a negative tolerance) drives dt down by 10 per restart until dt / 10 is
    a change DDFT::change_timestep ignores (< 1e-10, src/DDFT.cpp:184-186); the reference's restart loop would then
    repeat forever.  Both sides stop there with the density of the step's start."""
    shape, dx = (12, 12, 12), 0.1
    rho0 = density(shape, 0.5, 23, 0.1)
    w = oracle.fmt_weights(shape, dx, 1.0)
    o = oracle.DFT(shape, dx, 1.0, mf=True, kT=1.0)
    s = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mean_field=True, kT=1.0)
    do = oracle.DDFT(o, rho0, dt=1e-4, dt_max=1e-4, tol=-1.0)
    dg = cdft.DDFT(s, rho0, dt=1e-4, dt_max=1e-4, tol=-1.0)
    with pytest.raises(RuntimeError):
        do.step()
    with pytest.raises(cdft.cdft.DdftStalled):
        dg.step()
    assert np.abs(dg.density() - rho0).max() <= 1e-15 * rho0.max()
    assert np.abs(do.density() - rho0).max() <= 1e-15 * rho0.max()
