"""Open-boundary DDFT on the device (DDFT_IF_Open, mduran/src/DDFT_IF_Open.cpp:25-463) against the oracle's
restatement (pinned in tests/test_oracle_open_ddft.py): the GPU DST-I against scipy's, step-by-step replay
of F, dt, the restart and evaluation counts and the density (1e-10), the boundary held at the background,
and an ideal-gas Dirichlet mode decaying at its discrete eigenvalue."""
import numpy as np
import pytest
from scipy import fft as sfft

import oracle

pytestmark = pytest.mark.gpu


def density(shape, rho_b, seed, amp):
    u = oracle.splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    return rho_b * (1.0 + amp * (2.0 * u - 1.0))


def close(a, b, tol):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("shape", [(7, 11, 15), (15, 31, 63), (95, 47, 23)])
def test_dst1_matches_scipy(cdft, shape):
    a = np.random.default_rng(2).standard_normal(shape)
    b = cdft.dst1_3d(a)
    ref = sfft.dstn(a, type=1)
    assert np.max(np.abs(b - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("shape,steps", [((16, 16, 16), 3), ((24, 32, 48), 2)])
def test_open_ddft_replays_the_oracle(cdft, shape, steps):
    dx, hsd, bg = 0.1, 1.0, 0.5
    rho0 = density(shape, bg, 41, 0.1)
    w = oracle.fmt_weights(shape, dx, hsd)
    o = oracle.DFT(shape, dx, hsd, mf=True, kT=1.0)
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mean_field=True, kT=1.0)
    do, dg = oracle.OpenDDFT(o, rho0, bg), cdft.OpenDDFT(s, rho0, bg)
    for i in range(steps):
        a, b = do.step(), dg.step()
        assert (a["calls"], a["restarts"], a["evaluations"]) == (b["calls"], b["restarts"], b["evaluations"]), (a, b)
        assert close(a["F"], b["F"], 1e-10), (i, a["F"], b["F"])
        assert close(a["dt"], b["dt"], 1e-14) and close(a["time"], b["time"], 1e-14)
    r0, r1 = do.density(), dg.density()
    assert np.abs(r0 - r1).max() <= 1e-10 * np.abs(r0).max()
    bnd = np.zeros(shape, dtype=bool)
    bnd[0], bnd[:, 0], bnd[:, :, 0] = True, True, True
    assert np.all(r1[bnd] == bg)
    assert b["rhs_max"] > 0


def test_ideal_gas_dirichlet_mode(cdft):
    shape, dx, bg, eps = (12, 16, 18), 0.1, 0.5, 1e-7
    ix, iy, iz = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    mode = np.sin(np.pi * ix / shape[0]) * np.sin(np.pi * iy / shape[1]) * np.sin(np.pi * iz / shape[2])
    s = cdft.System(shape, dx, 1.0, model=cdft.MODEL_NONE, ideal=True)
    d = cdft.OpenDDFT(s, bg + eps * mode, bg, dt=1e-4, dt_max=1e-4)
    for _ in range(4):
        info = d.step()
    rho = d.density()
    lam = sum(2 / dx ** 2 * (np.cos(np.pi / n) - 1) for n in shape)
    amp = np.sum((rho - bg) * mode) / np.sum(mode * mode)
    assert abs(amp / (eps * np.exp(lam * info["time"])) - 1) < 1e-5

