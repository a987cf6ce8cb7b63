"""Pins of the oracle's DFT terms, mean field, FIRE2 and DDFT (CPU only).

* Mean field: brute-force convolution of the real-space kernel
  (Interaction_Base::checkCalc recipe, src/Interaction.cpp:536-563) and the
  uniform-density closed form F_mf / V = a_vdw rho^2 / 2 (include/Interaction.h:53).
* a_vdw converges to the continuum (2 pi / kT) int r^2 w_att(r) dr as dx -> 0
  (Potential1::getVDW_Parameter, include/Potential1.h:85-90), sanity only.
* Ideal gas and external field: closed forms (src/DFT.cpp:388-394,
  src/Species.cpp:51-62).
* DDFT: mass conserved to 1e-12; a small ideal-gas Fourier mode decays at the
  discrete diffusion rate Lambda_k (src/DDFT.cpp:70-92).
* FIRE2: F decreases over a run and the raw gradient shrinks.
"""
import numpy as np
from scipy import integrate

import oracle


def seeded(shape, rho_b, seed, amp):
    u = oracle.splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    return rho_b * (1 + amp * (2 * u - 1))


def test_mean_field_brute_force():
    shape, dx, kT = (10, 10, 10), 0.25, 0.8
    rho = seeded(shape, 0.5, 4, 0.3)
    w, a = oracle.mf_kernel(shape, dx, kT=kT)
    o = oracle.DFT(shape, dx, 1.0, fmt=False, ideal=False, mf=True, kT=kT)
    parts, dF = o.eval(rho)
    dV = dx ** 3
    idx = np.indices(shape).reshape(3, -1).T
    diff = (idx[:, None, :] - idx[None, :, :]) % np.array(shape)
    flat = (diff[..., 0] * shape[1] + diff[..., 1]) * shape[2] + diff[..., 2]
    v = dV * (w.reshape(-1)[flat] @ rho.reshape(-1))  # v_i = dV sum_j w(i-j) rho_j
    assert np.abs(v - dF.reshape(-1)).max() <= 1e-12 * np.abs(v).max()
    assert abs(parts[3] - 0.5 * rho.reshape(-1) @ v) <= 1e-12 * abs(parts[3])
    assert abs(parts[4]) == 0.0  # mu = 0, no V_ext


def test_mean_field_uniform_and_continuum():
    kT = 1.0
    for dx, N in [(0.1, 64), (0.05, 128)]:
        shape = (N, N, N)
        o = oracle.DFT(shape, dx, 1.0, fmt=False, ideal=False, mf=True, kT=kT)
        rb = 0.6
        parts, dF = o.eval(np.full(shape, rb))
        V = N ** 3 * dx ** 3
        assert abs(parts[3] / V - 0.5 * o.a_vdw * rb ** 2) <= 1e-12 * abs(parts[3] / V)
    # continuum: int w_att d^3r = 4 pi int_0^rc r^2 Watt(r) dr with the WCA
    # continuation Vmin inside rmin (twice Potential1::getVDW_Parameter's 2 pi form,
    # since F_mf/V = a_vdw rho^2 / 2)
    rc, rmin = 2.5, 2 ** (1 / 6)
    shift = 4 * (rc ** -12 - rc ** -6)
    vmin = 4 * (rmin ** -12 - rmin ** -6) - shift
    cont = 4 * np.pi * (vmin * rmin ** 3 / 3 + integrate.quad(lambda r: r * r * (4 * (r ** -12 - r ** -6) - shift), rmin, rc)[0])
    assert abs(o.a_vdw / cont - 1) < 5e-3


def test_ideal_and_external_closed_forms():
    shape, dx = (8, 8, 8), 0.2
    rho = seeded(shape, 0.3, 2, 0.5)
    vext = seeded(shape, 1.0, 3, 0.9)
    mu = -0.7
    o = oracle.DFT(shape, dx, 1.0, fmt=False, ideal=True, mu=mu, vext=vext)
    parts, dF = o.eval(rho)
    dV = dx ** 3
    assert abs(parts[1] - dV * np.sum(rho * np.log(rho) - rho)) <= 1e-13 * abs(parts[1])
    assert abs(parts[4] - (dV * np.sum(vext * rho) - mu * dV * rho.sum())) <= 1e-13 * abs(parts[4])
    assert np.abs(dF - dV * (np.log(rho) + vext - mu)).max() <= 1e-15


def test_ddft_mass_and_mode_decay():
    shape, dx = (16, 16, 16), 0.1
    x = np.arange(16)
    amp = 1e-7
    rho0 = 0.5 * (1 + amp * np.cos(2 * np.pi * 2 * x / 16))[:, None, None] * np.ones(shape)
    o = oracle.DFT(shape, dx, 1.0, fmt=False, ideal=True)
    d = oracle.DDFT(o, rho0, dt=1e-4, dt_max=1e-4)
    for _ in range(5):
        r = d.step()
    rho = d.density()
    lam = 2 / dx ** 2 * (np.cos(2 * np.pi * 2 / 16) - 1)
    c_meas = (rho[:, 0, 0] - 0.5) @ np.cos(2 * np.pi * 2 * x / 16) / 8
    assert abs(c_meas / (0.5 * amp * np.exp(lam * r["time"])) - 1) < 1e-5
    assert abs(rho.sum() / rho0.sum() - 1) < 1e-12
    # with FMT + mean field the mass is conserved too
    rho1 = seeded(shape, 0.5, 8, 0.1)
    o2 = oracle.DFT(shape, dx, 1.0, mf=True, kT=1.0)
    d2 = oracle.DDFT(o2, rho1)
    for _ in range(3):
        d2.step()
    assert abs(d2.density().sum() / rho1.sum() - 1) < 1e-12


def test_fire_decreases_free_energy():
    shape, dx = (12, 12, 12), 0.1
    rho0 = seeded(shape, 0.4, 5, 0.2)
    o = oracle.DFT(shape, dx, 1.0, mu=-0.5)
    F0 = o.eval(rho0)[0][0]
    f = oracle.Fire(o, rho0)
    for _ in range(200):
        rc, info = f.step()
    assert info["F"] < F0
    assert info["fmax_raw"] < 1.0
