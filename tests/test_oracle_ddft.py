"""Pin of the oracle's DDFT integrating-factor step (src/DDFT.cpp:113-292) by linear response.

Ideal gas plus a mean field with kernel w at uniform density rho0: a small density mode delta rho ~ cos(k x)
obeys d(delta rho_k)/dt = Lambda_k (1 + rho0 w_k) delta rho_k, with Lambda_k = 2 (cos(k dx) - 1) / dx^2 the
eigenvalue of the discrete Laplacian that g . x reduces to at uniform density (DDFT::g_dot_x, :320-351) and w_k the
discrete transform of the kernel.  The integrating factor is exact for the Lambda_k part; the excess part enters
through U1 R0 + U2 (R1 - R0) (apply_diffusion_propogator, :239-292) with R1 from the implicit fixed point (:141-165).
Converged fixed point: the scheme is second order, the error drops 4x per halving of dt.  One fixed-point pass
(R1 = R0 at the old density): first order.  A wrong U1, U2 or g . x breaks either the rate or the order.
"""
import numpy as np
import pytest

import oracle

SHAPE, DX, RHO0, EPS, T = (16, 8, 8), 0.1, 0.5, 1e-6, 0.04


def _kernel():
    def minimg(n):
        i = np.arange(n)
        return np.minimum(i, n - i) * DX
    X, Y, Z = np.meshgrid(minimg(SHAPE[0]), minimg(SHAPE[1]), minimg(SHAPE[2]), indexing="ij")
    return -0.0228 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / (2 * 0.15 ** 2))


def mode_setup(m=1):
    w = _kernel()
    wk = np.fft.fftn(w).real[m, 0, 0]
    k = 2 * np.pi * m / (SHAPE[0] * DX)
    rate = 2 * (np.cos(k * DX) - 1) / DX ** 2 * (1 + RHO0 * wk)
    x = np.arange(SHAPE[0])
    mode = np.cos(2 * np.pi * m * x / SHAPE[0])[:, None, None] * np.ones(SHAPE)
    return w, rate, mode


def amplitude(rho, mode):
    return float(((rho - RHO0) * mode).sum() / (mode * mode).sum() / RHO0)


def oracle_run(dt, tol):
    w, rate, mode = mode_setup()
    d = oracle.DFT(SHAPE, DX, 1.0, fmt=False, ideal=True, mf=True)
    d.set_mean_field(w, float(w.sum() * DX ** 3))
    dd = oracle.DDFT(d, RHO0 * (1 + EPS * mode), dt=dt, dt_max=dt, tol=tol)
    for _ in range(int(round(T / dt))):
        info = dd.step()
    return amplitude(dd.density(), mode) / (EPS * np.exp(rate * info["time"])) - 1, info


def test_mode_decays_at_the_linear_response_rate_to_second_order():
    w, rate, _ = mode_setup()
    assert -0.6 < RHO0 * np.fft.fftn(w).real[1, 0, 0] < -0.4  # a real excess term, still stable
    errs = []
    for dt in (4e-3, 2e-3, 1e-3):
        e, info = oracle_run(dt, 1e-14)
        assert info["restarts"] == 0 and abs(info["time"] - T) < 1e-12
        errs.append(e)
    assert abs(errs[0]) < 5e-5
    for a, b in zip(errs, errs[1:]):
        assert 3.6 < a / b < 4.4


def test_one_fixed_point_pass_is_first_order():
    errs = [oracle_run(dt, 0.999)[0] for dt in (4e-3, 2e-3)]
    assert abs(errs[0]) > 1e-3
    assert 1.8 < errs[0] / errs[1] < 2.2
