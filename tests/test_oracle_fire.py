"""Pins of the oracle's FIRE2 minimizer (fireMinimizer2, src/Minimizer.cpp:117-327) by what the mathematics fixes.

* The ideal gas in an external field is separable: its minimum is rho_i = exp(mu - V_i) exactly, and FIRE run to
  a tight tolerance must land there.
* The alias map rho = rho_min + c y^2 / (1 + y^2) (src/FMT_Species.cpp:29-93) and its gradient factor
  2 c y / (1 + y^2)^2 are checked against central differences of F(rho(y)).
* The trajectory: a transcription, written here from the published algorithm (FIRE 2.0 with semi-implicit Euler,
  Guenole et al., Comput. Mater. Sci. 175, 109584 (2020), Algorithm 2) with the reference's constants
  (src/Minimizer.cpp:117-147, include/Minimizer.h:111-146) and the reference's three additions to it (velocity
  scaled by 0.1 instead of zeroed after an uphill step, :218; the dt_max control after 10 uphill or backtracked
  steps, :245-250; dt_best), must follow the oracle step for step on the separable ideal-gas problem, where F and
  its gradient are closed forms.  A changed constant (f_inc, f_dec, f_alpha, alpha_start, N_delay), a different
  mixing rule or a wrong alias factor changes the trajectory at once.
"""
import numpy as np
import pytest

import oracle

SHAPE, DX, HSD, MU = (6, 6, 6), 0.1, 1.0, -1.0
RHO_MIN = 1e-18  # SMALL_VALUE


def _problem(seed=3):
    rng = np.random.default_rng(seed)
    x = np.arange(SHAPE[0]) * DX * 2 * np.pi / (SHAPE[0] * DX)
    vext = 0.8 * np.cos(x)[:, None, None] + 0.3 * rng.uniform(-1, 1, SHAPE)
    rho0 = np.exp(MU) * (1 + 0.4 * rng.uniform(-1, 1, SHAPE))
    return vext, rho0


def _alias_c():
    dV = DX ** 3
    return (1.0 - RHO_MIN * 4 * np.pi * HSD ** 3 / 3) / dV


def _rho_of(y):
    c = _alias_c()
    return RHO_MIN + c * y * y / (1 + y * y)


def _F_dF(rho, vext):
    dV = DX ** 3
    F = dV * np.sum(rho * np.log(rho) - rho) + dV * np.sum(vext * rho) - MU * dV * np.sum(rho)
    return F, dV * (np.log(rho) + vext - MU)


def _transcribed_fire(y, vext, dt, dt_max, steps):
    """FIRE 2.0 (semi-implicit Euler) on the alias coordinates y, returning per-step (F, dt, alpha) and the state."""
    c = _alias_c()
    dt_min, n_delay, a_start, f_dec, f_inc, f_a, n_neg_max = 0.0, 5, 0.1, 0.5, 1.1, 0.99, 20

    def grad(y):
        F, dF = _F_dF(_rho_of(y), vext)
        return F, dF * 2 * c * y / (1 + y * y) ** 2

    F, g = grad(y)
    v = np.zeros_like(y)
    alpha, n_pos, n_neg = a_start, 0, 0
    dt_best, backtracks = dt, 0
    out = []
    for it in range(1, steps + 1):
        P = -np.dot(v.ravel(), g.ravel())  # the force is -g
        if P > 0:
            n_pos, n_neg = n_pos + 1, 0
            if n_pos > n_delay:
                dt = min(dt * f_inc, dt_max)
                alpha *= f_a
        else:
            n_pos, n_neg = 0, n_neg + 1
            assert n_neg <= n_neg_max
            if not it < n_delay:  # initial delay
                if dt * f_dec >= dt_min:
                    dt *= f_dec
                alpha = a_start
            y = y - 0.5 * dt * v
            v = 0.1 * v
            backtracks += 1
        v = v - dt * g
        vn, fn = np.sqrt(np.sum(v * v)), np.sqrt(np.sum(g * g))
        v = (1 - alpha) * v - alpha * (vn / fn) * g
        y = y + dt * v
        F, g = grad(y)
        dt_best = max(dt, dt_best)
        if backtracks >= 10:
            dt_max = min(dt_best, 0.9 * dt_max)
            dt_best, backtracks = 0.0, 0
        out.append((F, dt, alpha))
    return out, y, v


def test_alias_map_gradient_is_the_chain_rule():
    """dF/dy = dF/drho * 2 c y / (1 + y^2)^2, by central differences of F(rho(y)) in every component."""
    vext, rho0 = _problem()
    c = _alias_c()
    y = np.sqrt(np.abs(((rho0 - RHO_MIN) / c) / (1 - (rho0 - RHO_MIN) / c)))
    assert np.allclose(_rho_of(y), rho0, rtol=1e-13, atol=0)
    d = oracle.DFT(SHAPE, DX, HSD, fmt=False, ideal=True, mu=MU, vext=vext)
    f = oracle.Fire(d, rho0, 1e-3, 1e-2)
    rho, x, v, dF_raw = f.state()
    assert np.allclose(x, y, rtol=1e-13, atol=0)
    _, g_exact = _F_dF(rho0, vext)
    g_alias = g_exact * 2 * c * y / (1 + y * y) ** 2
    h = 1e-6
    for i in [(0, 0, 0), (1, 2, 3), (5, 4, 1)]:
        yp, ym = y.copy(), y.copy()
        yp[i] += h
        ym[i] -= h
        fd = (_F_dF(_rho_of(yp), vext)[0] - _F_dF(_rho_of(ym), vext)[0]) / (2 * h)
        assert abs(fd - g_alias[i]) <= 1e-7 * abs(g_alias[i])
    # the oracle's first step moves x by dt * (mixed velocity) built from this gradient
    rc, info = f.step()
    _, x1, _, _ = f.state()
    dt, a = info["dt"], info["alpha"]
    v1 = -dt * g_alias
    v1 = (1 - a) * v1 - a * np.sqrt(np.sum(v1 ** 2) / np.sum(g_alias ** 2)) * g_alias
    assert np.max(np.abs(x1 - (y + dt * v1))) <= 1e-13 * np.max(np.abs(y))


def test_fire_follows_the_published_algorithm():
    vext, rho0 = _problem()
    c = _alias_c()
    y0 = np.sqrt(np.abs(((rho0 - RHO_MIN) / c) / (1 - (rho0 - RHO_MIN) / c)))
    d = oracle.DFT(SHAPE, DX, HSD, fmt=False, ideal=True, mu=MU, vext=vext)
    dt0, dtmax0 = 2e-3, 2e-2
    f = oracle.Fire(d, rho0, dt0, dtmax0)
    ref, yK, vK = _transcribed_fire(y0, vext, dt0, dtmax0, 120)
    ups = 0
    for k, (F, dt, alpha) in enumerate(ref):
        rc, info = f.step()
        assert rc == 0
        assert abs(info["F"] - F) <= 1e-12 * abs(F), k
        assert info["dt"] == pytest.approx(dt, rel=1e-14), k
        assert info["alpha"] == pytest.approx(alpha, rel=1e-14), k
        ups += alpha == 0.1 and k > 5
    _, x, v, _ = f.state()
    assert np.max(np.abs(x - yK)) <= 1e-11 * np.max(np.abs(yK))
    assert np.max(np.abs(v - vK)) <= 1e-9 * np.max(np.abs(vK))
    assert ups > 0  # the trajectory exercises the uphill branch


def test_fire_reaches_the_ideal_gas_equilibrium():
    """FIRE run until max |dF/drho| / dV <= 1e-11 lands on rho = exp(mu - V_ext) to that accuracy."""
    vext, rho0 = _problem(5)
    d = oracle.DFT(SHAPE, DX, HSD, fmt=False, ideal=True, mu=MU, vext=vext)
    f = oracle.Fire(d, rho0, 2e-3, 2e-2)
    for k in range(20000):
        rc, info = f.step()
        assert rc == 0
        if info["fmax_raw"] <= 1e-11:
            break
    assert info["fmax_raw"] <= 1e-11
    rho, _, _, _ = f.state()
    exact = np.exp(MU - vext)
    assert np.max(np.abs(rho / exact - 1)) <= 2e-11
