"""Pins of the oracle's cell-integrated FMT weights (src/FMT_Species.cpp:112-496).

* Sum rules of the trilinear-hat partition of unity (fillUniform,
  include/Fundamental_Measures.h:20-31): sum w_eta = pi d^3/6, sum w_s = pi d^2,
  sum w_v = 0, sum w_Tii = pi d^2/3, off-diagonal T sums = 0.
* Parity: scalar and tensor weights even, vector weights odd.
* Independent adaptive quadrature of single boundary cells (scipy) for the
  volume weight; cell-by-cell identities for the others: w_s = d w_eta / dR
  and w_s = tr w_T; exact first moments (the trilinear hat reproduces linear
  functions): sum_S X_S w_vx(S) = 4 pi R^3 / 3, sum_S X_S Y_S w_Txy(S) = 4 pi R^4 / 15.
* The reference-faithful double evaluation deviates from the accurate one by
  its documented roundoff only (asin near +-1): this records the size of the
  reference's own weight noise that makes shared weights necessary for 1e-11
  FMT parity.
"""
import warnings

import numpy as np
import pytest
from scipy import integrate

import oracle


def sums(w):
    return w.sum(axis=(1, 2, 3))


@pytest.mark.parametrize("dx,N,hsd", [(0.1, 32, 1.0), (0.05, 48, 1.0), (0.1, 24, 1.37), (0.04, 48, 1.0)])
def test_sum_rules_accurate(dx, N, hsd):
    w = oracle.fmt_weights((N, N, N), dx, hsd, mode=oracle.WEIGHTS_ACCURATE)
    s = sums(w)
    assert abs(s[0] / (np.pi * hsd ** 3 / 6) - 1) < 1e-13
    assert abs(s[1] / (np.pi * hsd ** 2) - 1) < 1e-13
    for a in (5, 8, 10):
        assert abs(s[a] / (np.pi * hsd ** 2 / 3) - 1) < 1e-13
    for a in (2, 3, 4, 6, 7, 9):
        assert abs(s[a]) < 1e-15


def test_sum_rules_anisotropic():
    dxs = (0.1, 0.08, 0.125)
    w = oracle.fmt_weights((24, 32, 20), dxs, 1.0, mode=oracle.WEIGHTS_ACCURATE)
    s = sums(w)
    assert abs(s[0] / (np.pi / 6) - 1) < 1e-13
    assert abs(s[1] / np.pi - 1) < 1e-13
    for a in (5, 8, 10):
        assert abs(s[a] / (np.pi / 3) - 1) < 1e-13


def test_parity():
    w = oracle.fmt_weights((32, 32, 32), 0.1, 1.0, mode=oracle.WEIGHTS_ACCURATE)

    def reflect(a):
        return np.roll(a[::-1, ::-1, ::-1], 1, axis=(0, 1, 2))

    for a in range(11):
        sign = -1 if a in (2, 3, 4) else 1
        assert np.abs(w[a] - sign * reflect(w[a])).max() < 1e-18
    # w_vx is odd in x and even in y, z
    vx = w[2]
    assert np.abs(vx + np.roll(vx[::-1], 1, axis=0)).max() < 1e-18
    assert np.abs(vx - np.roll(vx[:, ::-1], 1, axis=1)).max() < 1e-18


def hat(u, h):
    return np.clip(1 - np.abs(u) / h, 0, None)


def eta_cell_quadrature(S, dx, R):
    """Direct adaptive quadrature of w_eta(S) = int_sphere hat(x - S dx)."""
    cx, cy, cz = (s * dx for s in S)
    total = 0.0
    # integrate z analytically over the chord, x and y adaptively per sub-square
    for x0 in (cx - dx, cx):
        for y0 in (cy - dx, cy):
            def inner(y, x):
                h2 = R * R - x * x - y * y
                if h2 <= 0:
                    return 0.0
                h = np.sqrt(h2)
                lo, hi = max(cz - dx, -h), min(cz + dx, h)
                if hi <= lo:
                    return 0.0
                # integral of (1 - |z - cz|/dx) over [lo, hi], split at cz
                def piece(a, b, sgn):
                    if b <= a:
                        return 0.0
                    return (b - a) - sgn * ((b - cz) ** 2 - (a - cz) ** 2) / (2 * dx)
                val = piece(lo, min(hi, cz), -1) + piece(max(lo, cz), hi, +1)
                return hat(x - cx, dx) * hat(y - cy, dx) * val
            v, _ = integrate.dblquad(inner, x0, x0 + dx, y0, y0 + dx, epsabs=1e-14, epsrel=1e-11)
            total += v
    return total


@pytest.mark.parametrize("S", [(4, 2, 1), (3, 3, 2), (5, 0, 0), (1, 4, 2)])
@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
def test_boundary_cells_against_quadrature(S):
    dx, R, N = 0.1, 0.5, 16
    w = oracle.fmt_weights((N, N, N), dx, 2 * R, mode=oracle.WEIGHTS_ACCURATE)
    idx = tuple(s % N for s in S)
    q = eta_cell_quadrature(S, dx, R)
    # the quadrature itself is limited by kinks of the chord length (~1e-9 dV)
    assert abs(w[0][idx] - q) < 5e-9 * dx ** 3


def test_faithful_weights_noise_is_bounded():
    """Reference-faithful double weights = accurate weights + asin roundoff noise."""
    for dx, N, bound in [(0.1, 32, 5e-6), (0.05, 48, 2e-4)]:
        wf = oracle.fmt_weights((N, N, N), dx, 1.0, mode=oracle.WEIGHTS_FAITHFUL)
        wa = oracle.fmt_weights((N, N, N), dx, 1.0, mode=oracle.WEIGHTS_ACCURATE)
        for a in range(11):
            assert np.abs(wf[a] - wa[a]).max() <= bound * np.abs(wa[a]).max()
        # and the noise is large enough to matter for 1e-11 FMT parity
        assert np.abs(wf[0] - wa[0]).max() > 1e-12 * np.abs(wa[0]).max()


def test_cell_identities():
    dx, N = 0.1, 32
    R = 0.4873  # a radius through no lattice point keeps w_eta(R) smooth
    h = 1e-5
    wp = oracle.fmt_weights((N, N, N), dx, 2 * (R + h), mode=oracle.WEIGHTS_ACCURATE)
    wm = oracle.fmt_weights((N, N, N), dx, 2 * (R - h), mode=oracle.WEIGHTS_ACCURATE)
    w = oracle.fmt_weights((N, N, N), dx, 2 * R, mode=oracle.WEIGHTS_ACCURATE)
    ds = (wp[0] - wm[0]) / (2 * h)
    assert np.abs(ds - w[1]).max() < 1e-7 * np.abs(w[1]).max()
    assert np.abs(w[5] + w[8] + w[10] - w[1]).max() < 1e-13 * np.abs(w[1]).max()


@pytest.mark.parametrize("dx,N,hsd", [(0.1, 32, 1.0), (0.07, 40, 1.3)])
def test_first_moments(dx, N, hsd):
    R = hsd / 2
    w = oracle.fmt_weights((N, N, N), dx, hsd, mode=oracle.WEIGHTS_ACCURATE)
    i = np.arange(N)
    X = np.where(i < N // 2, i, i - N) * dx  # minimal-image coordinates
    Xg, Yg, Zg = np.meshgrid(X, X, X, indexing="ij")
    assert abs((Xg * w[2]).sum() / (4 * np.pi * R ** 3 / 3) - 1) < 1e-12
    assert abs((Yg * w[3]).sum() / (4 * np.pi * R ** 3 / 3) - 1) < 1e-12
    assert abs((Xg * Yg * w[6]).sum() / (4 * np.pi * R ** 4 / 15) - 1) < 1e-12
    assert abs((Yg * Zg * w[9]).sum() / (4 * np.pi * R ** 4 / 15) - 1) < 1e-12
    # and the linear moment of the volume weight vanishes
    assert abs((Xg * w[0]).sum()) < 1e-16
