"""Pins the oracle's Asakura-Oosawa polymer (FMT_AO_Species: include/Species.h:380-461,
src/FMT_Species.cpp:647-738, src/FMT.cpp:393-422) before it checks the device.

F_AO's derivative is pinned by central differences of F for every model; the
uniform-density limit by the closed form of F_AO (psi constant); the reference's
own force, which stores one of the two tensor off-diagonal entries
(include/Species.h:406-425), is shown to differ from the derivative exactly by
those entries (it agrees for the scalar-vector models Rosenfeld and RSLT)."""
import numpy as np
import pytest

import oracle

SHAPE, DX, HSD, RP, RHO_P = (12, 10, 14), 0.1, 1.0, 0.3, 0.4


def density(seed, rho0=0.45):
    """Structure along all three axes and their diagonals, so that the off-diagonal tensor measures matter."""
    rng = np.random.default_rng(seed)
    x = [2 * np.pi * g / n for g, n in zip(np.meshgrid(*[np.arange(n) for n in SHAPE], indexing="ij"), SHAPE)]
    return rho0 * (1 + 0.25 * np.cos(x[0]) * np.cos(x[1]) + 0.2 * np.sin(x[1] + x[2]) + 0.15 * np.cos(x[0] - x[2])
                   + 0.05 * (2 * rng.random(SHAPE) - 1))


@pytest.mark.parametrize("model", [oracle.ROSENFELD, oracle.RSLT, oracle.ESFMT, oracle.WBI, oracle.WBII])
def test_ao_force_is_the_derivative_of_F(model):
    rho = density(30 + model)
    f = oracle.FMT(model, SHAPE, DX, HSD, A=1.0, B=-0.5)
    f.set_ao(RP, RHO_P)
    F, dF = f.eval(rho)
    rng = np.random.default_rng(40 + model)
    dV = DX ** 3
    for _ in range(2):
        v = 2 * rng.random(SHAPE) - 1
        h = 1e-5
        fd = (f.eval(rho + h * v)[0] - f.eval(rho - h * v)[0]) / (2 * h)
        assert abs(fd - np.sum(dF * v)) <= 2e-8 * np.sum(np.abs(dF * v)), (fd, np.sum(dF * v))
    # the reference's tensor count: identical without tensor; with it, its force is not the derivative of F
    g = oracle.FMT(model, SHAPE, DX, HSD, A=1.0, B=-0.5)
    g.set_ao(RP, RHO_P, exact=False)
    Fr, dFr = g.eval(rho)
    assert Fr == F
    if model in (oracle.ROSENFELD, oracle.RSLT):
        assert np.array_equal(dFr, dF)
    else:
        v = 2 * rng.random(SHAPE) - 1
        h = 1e-5
        fd = (f.eval(rho + h * v)[0] - f.eval(rho - h * v)[0]) / (2 * h)
        err_exact = abs(fd - np.sum(dF * v))
        err_ref = abs(fd - np.sum(dFr * v))
        assert err_ref > 100 * err_exact, (err_ref, err_exact)


def test_ao_uniform_density_closed_form():
    """Uniform rho: psi = sum_a dPhi^p_a W^p_a(k=0) is a constant, so F_AO = rho_p N (1 - exp(-psi)); the AO
    force is uniform too."""
    rho = np.full(SHAPE, 0.5)
    f0 = oracle.FMT(oracle.WBI, SHAPE, DX, HSD)
    F0, d0 = f0.eval(rho)
    f = oracle.FMT(oracle.WBI, SHAPE, DX, HSD)
    f.set_ao(RP, RHO_P)
    F, dF = f.eval(rho)
    N, dV = np.prod(SHAPE), DX ** 3
    psi = -np.log(1 - (F - F0) / (RHO_P * N * dV))
    assert psi > 0
    assert np.max(np.abs(dF - dF.mean())) <= 1e-12 * abs(dF.mean())
    # psi = rho_p-independent: doubling rho_p doubles F_AO
    f2 = oracle.FMT(oracle.WBI, SHAPE, DX, HSD)
    f2.set_ao(RP, 2 * RHO_P)
    F2, _ = f2.eval(rho)
    assert abs((F2 - F0) - 2 * (F - F0)) <= 1e-12 * abs(F - F0)
