"""Pins of the oracle's FMT F + dF (src/FMT.cpp:310-427).

* Brute force at 8^3: weighted densities by direct-space summation
  (the getBruteForceWeightedDensity recipe, include/Species.h:249-265),
  Phi and its 19 derivatives site by site, and dF_i = dV sum_K sum_j
  dPhi_j(K) dm_j(K)/drho_i with the true (correlation) kernel -- this pins
  the FFT convolution conventions and the vector parity sign independently.
* Central finite differences of F reproduce dF.
* Uniform density: F/V = f_ex, dF = dV mu_ex (accurate weights).
* Translation covariance.
* Committed config-1 fixture guards regressions.
"""
import os

import numpy as np
import pytest

import oracle

TI = [[5, 6, 7], [6, 8, 9], [7, 9, 10]]


def seeded(shape, rho_b, seed, amp):
    u = oracle.splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    return rho_b * (1 + amp * (2 * u - 1))


def brute_force(model, rho, dx, hsd, w, A=1.0, B=0.0):
    shape = rho.shape
    N = rho.size
    dV = dx ** 3
    idx = np.indices(shape).reshape(3, -1).T
    # K - J with periodic wrap for all pairs
    diff = (idx[:, None, :] - idx[None, :, :]) % np.array(shape)
    flat = (diff[..., 0] * shape[1] + diff[..., 1]) * shape[2] + diff[..., 2]
    wf = w.reshape(11, -1)
    r = rho.reshape(-1)
    n = np.stack([wf[a][flat] @ r for a in range(11)])  # n_a(K)
    d = hsd
    F = 0.0
    G = np.zeros((11, N))
    for K in range(N):
        m = np.zeros(19)
        m[0] = n[0, K]
        m[1], m[2], m[3] = n[1, K] / d ** 2, n[1, K] / d, n[1, K]
        for k in range(3):
            m[4 + k] = n[2 + k, K] / d
            m[7 + k] = n[2 + k, K]
        tensor = model >= oracle.ESFMT
        for j in range(3):
            for k in range(3):
                m[10 + 3 * j + k] = n[TI[j][k], K] if tensor else 0.0
        phi, dphi = oracle.phi(model, m, A, B)
        F += phi
        G[0, K] += dphi[0]
        G[1, K] += dphi[1] / d ** 2 + dphi[2] / d + dphi[3]
        for k in range(3):
            G[2 + k, K] += dphi[4 + k] / d + dphi[7 + k]
        if tensor:
            for j in range(3):
                for k in range(3):
                    G[TI[j][k], K] += dphi[10 + 3 * j + k]
    # dF_i = dV sum_K sum_a G_a(K) w_a(K - i)
    dF = dV * sum(G[a] @ wf[a][flat] for a in range(11))
    return F * dV, dF.reshape(shape)


@pytest.mark.parametrize("model,A,B", [(oracle.ROSENFELD, 1, 0), (oracle.WBI, 1, -1), (oracle.ESFMT, 1.3, -0.4)])
def test_brute_force_8(model, A, B):
    shape, dx, hsd = (8, 8, 8), 0.2, 1.0
    rho = seeded(shape, 0.4, 99 + model, 0.3)
    w = oracle.fmt_weights(shape, dx, hsd, mode=oracle.WEIGHTS_ACCURATE)
    F0, dF0 = oracle.fmt_eval(model, rho, dx, hsd, weights=w, A=A, B=B)
    F1, dF1 = brute_force(model, rho, dx, hsd, w, A, B)
    assert abs(F0 - F1) <= 1e-12 * abs(F1)
    assert np.abs(dF0 - dF1).max() <= 1e-11 * np.abs(dF1).max()


def test_central_differences_of_F():
    shape, dx, hsd = (12, 12, 12), 0.1, 1.0
    rho = seeded(shape, 0.5, 3, 0.2)
    w = oracle.fmt_weights(shape, dx, hsd, mode=oracle.WEIGHTS_ACCURATE)
    h = oracle.FMT(oracle.WBI, shape, dx, hsd, weights=w)
    F, dF = h.eval(rho)
    rng = np.random.default_rng(0)
    for _ in range(5):
        i = tuple(rng.integers(0, 12, 3))
        eps = 1e-5 * rho[i]
        rp, rm = rho.copy(), rho.copy()
        rp[i] += eps
        rm[i] -= eps
        num = (h.eval(rp)[0] - h.eval(rm)[0]) / (2 * eps)
        assert abs(num - dF[i]) <= 1e-7 * abs(dF[i])


def test_uniform_density():
    shape, dx, hsd = (24, 24, 24), 0.1, 1.0
    w = oracle.fmt_weights(shape, dx, hsd, mode=oracle.WEIGHTS_ACCURATE)
    for model in (oracle.ROSENFELD, oracle.RSLT, oracle.WBI, oracle.WBII):
        for rb in (0.3, 0.8):
            F, dF = oracle.fmt_eval(model, np.full(shape, rb), dx, hsd, weights=w)
            fex, mu = oracle.bulk(model, rb, hsd)
            V = np.prod(shape) * dx ** 3
            assert abs(F / V - fex) <= 1e-12 * abs(fex)
            assert np.abs(dF / dx ** 3 - mu).max() <= 1e-12 * abs(mu)


def test_translation_covariance():
    shape, dx, hsd = (16, 16, 16), 0.1, 1.0
    rho = seeded(shape, 0.5, 17, 0.2)
    w = oracle.fmt_weights(shape, dx, hsd)
    F0, dF0 = oracle.fmt_eval(oracle.WBI, rho, dx, hsd, weights=w)
    sh = (3, -5, 7)
    F1, dF1 = oracle.fmt_eval(oracle.WBI, np.roll(rho, sh, axis=(0, 1, 2)), dx, hsd, weights=w)
    assert abs(F0 - F1) <= 1e-13 * abs(F0)
    assert np.abs(np.roll(dF0, sh, axis=(0, 1, 2)) - dF1).max() <= 1e-13 * np.abs(dF0).max()


def test_config1_fixture():
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "config1_wb_32.npz"))
    shape, dx, hsd = (32, 32, 32), 0.1, 1.0
    rho = seeded(shape, 0.6, 0x5EED0001, 0.05)
    assert np.array_equal(rho.reshape(-1)[:16], g["rho_head"])
    for name, mode in [("faithful", oracle.WEIGHTS_FAITHFUL), ("accurate", oracle.WEIGHTS_ACCURATE)]:
        w = oracle.fmt_weights(shape, dx, hsd, mode=mode)
        F, dF = oracle.fmt_eval(oracle.WBI, rho, dx, hsd, weights=w)
        assert abs(F - float(g["F_" + name])) <= 1e-13 * abs(F)
        assert np.abs(dF[:2, :2, :4] - g["dF_corner_" + name]).max() <= 1e-13 * np.abs(dF).max()
