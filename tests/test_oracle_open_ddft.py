"""Pins the oracle's open-boundary DDFT (DDFT_IF_Open, mduran/src/DDFT_IF_Open.cpp:25-463) before it
checks the device: its 3D DST-I against scipy's DST-I (the FFTW RODFT00 convention), the transform
applied twice = 8 Nx Ny Nz x identity, an ideal-gas Dirichlet mode that decays at the rate of its
discrete Laplacian eigenvalue, and the boundary held at the background."""
import numpy as np
import pytest
from scipy import fft as sfft

import oracle


def test_dst1_matches_scipy():
    a = np.random.default_rng(1).standard_normal((5, 7, 11))
    b = oracle.dst1_3d(a)
    ref = sfft.dstn(a, type=1)
    assert np.max(np.abs(b - ref)) <= 1e-13 * np.max(np.abs(ref))
    twice = oracle.dst1_3d(b)
    assert np.max(np.abs(twice / (8 * 6 * 8 * 12) - a)) <= 1e-13


def test_ideal_gas_dirichlet_mode_decays_at_its_eigenvalue():
    shape, dx, bg, eps = (12, 16, 18), 0.1, 0.5, 1e-7
    ix, iy, iz = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    mode = np.sin(np.pi * ix / shape[0]) * np.sin(np.pi * iy / shape[1]) * np.sin(np.pi * iz / shape[2])
    rho0 = bg + eps * mode
    o = oracle.DFT(shape, dx, 1.0, fmt=False, ideal=True)
    d = oracle.OpenDDFT(o, rho0, bg, dt=1e-4, dt_max=1e-4)
    for _ in range(4):
        info = d.step()
    rho = d.density()
    lam = sum(2 / dx ** 2 * (np.cos(np.pi / n) - 1) for n in shape)
    amp = np.sum((rho - bg) * mode) / np.sum(mode * mode)
    assert abs(amp / (eps * np.exp(lam * info["time"])) - 1) < 1e-5
    bnd = (ix == 0) | (iy == 0) | (iz == 0)
    assert np.all(rho[bnd] == bg)


def test_boundary_is_set_to_the_background():
    shape, dx = (12, 12, 12), 0.1
    rng = np.random.default_rng(3)
    rho0 = 0.4 * (1 + 0.05 * (2 * rng.random(shape) - 1))
    o = oracle.DFT(shape, dx, 1.0, mf=True, kT=1.0)
    d = oracle.OpenDDFT(o, rho0, 0.42)
    info = d.step()
    rho = d.density()
    assert np.all(rho[0] == 0.42) and np.all(rho[:, 0] == 0.42) and np.all(rho[:, :, 0] == 0.42)
    assert np.all(rho[1:, 1:, 1:] != 0.42) and info["calls"] == 1
