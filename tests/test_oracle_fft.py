"""Pins of the oracle's FFT (stand-in for FFTW's r2c/c2r, include/DFT_LinAlg.h:163-288).

Checks: committed numpy golden vectors, live numpy.fft, a naive O(N^2) DFT,
round trip = N * identity, Parseval, and Hermitian symmetry.
"""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 24, 30, 48, 64, 96, 128, 256])
def test_fft1d_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    assert np.abs(oracle.fft1d(x) - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())
    assert np.abs(oracle.fft1d(x, +1) - np.fft.ifft(x) * n).max() <= 1e-14 * max(1.0, np.abs(ref).max())


def naive_dft3(x):
    Nx, Ny, Nz = x.shape
    out = np.zeros((Nx, Ny, Nz // 2 + 1), dtype=complex)
    ix, iy, iz = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")
    for a in range(Nx):
        for b in range(Ny):
            for c in range(Nz // 2 + 1):
                ph = np.exp(-2j * np.pi * (a * ix / Nx + b * iy / Ny + c * iz / Nz))
                out[a, b, c] = (x * ph).sum()
    return out


def test_r2c_naive_dft():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 6, 4))
    assert np.abs(oracle.rfftn(x) - naive_dft3(x)).max() < 1e-12


@pytest.mark.parametrize("shape", [(8, 8, 8), (6, 12, 18), (32, 32, 32), (16, 24, 48)])
def test_r2c_c2r_numpy(shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    s = oracle.rfftn(x)
    ref = np.fft.rfftn(x)
    assert np.abs(s - ref).max() / np.abs(ref).max() < 1e-14
    back = oracle.irfftn(s, shape[2])
    assert np.abs(back - x * x.size).max() / (x.size * np.abs(x).max()) < 1e-14
    # Parseval on the full spectrum
    full = np.fft.fftn(x)
    assert abs((np.abs(full) ** 2).sum() / x.size - (x ** 2).sum()) / (x ** 2).sum() < 1e-12


def test_golden_vectors():
    d = np.load(os.path.join(GOLDEN, "fft_vectors.npz"))
    for key in d.files:
        if not key.startswith("in_"):
            continue
        tag = key[3:]
        x = d["in_" + tag]
        assert np.abs(oracle.rfftn(x) - d["r2c_" + tag]).max() / np.abs(d["r2c_" + tag]).max() < 1e-14
