"""Pins the oracle's Hessian-vector product (restating FMT::calculate_d2Phi_dot_V and
FMT::add_second_derivative, src/FMT.cpp:181-280, 437-567, and DFT::matrix_dot_v_intern,
src/DFT.cpp:435-488) against central differences of the oracle's own dF, model by
model with the FMT part isolated, and by the symmetry u.(Hv) = v.(Hu)."""
import numpy as np
import pytest

import oracle

SHAPE, DX = (12, 10, 14), 0.1


def inputs(seed, rho0=0.7):
    rng = np.random.default_rng(seed)
    x = np.meshgrid(*[np.arange(n) for n in SHAPE], indexing="ij")
    wave = np.cos(2 * np.pi * x[0] / SHAPE[0]) * np.sin(4 * np.pi * x[2] / SHAPE[2])
    rho = rho0 * (1 + 0.3 * wave + 0.1 * (2 * rng.random(SHAPE) - 1))
    return rho, 2 * rng.random(SHAPE) - 1, 2 * rng.random(SHAPE) - 1


@pytest.mark.parametrize("model", [oracle.ROSENFELD, oracle.RSLT, oracle.ESFMT, oracle.WBI, oracle.WBII])
def test_fmt_hessian_matches_central_differences(model):
    rho, v, u = inputs(10 + model)
    d = oracle.DFT(SHAPE, DX, 1.0, model=model, A=1.0, B=-0.5, ideal=False)
    Hv = d.hess(rho, v)
    h = 1e-5
    fd = (d.eval(rho + h * v)[1] - d.eval(rho - h * v)[1]) / (2 * h)
    assert np.max(np.abs(fd - Hv)) <= 1e-8 * np.max(np.abs(Hv))
    Hu = d.hess(rho, u)
    assert abs(np.sum(u * Hv) - np.sum(v * Hu)) <= 1e-13 * np.sum(np.abs(u * Hv))


def test_dft_hessian_ideal_and_mean_field():
    rho, v, _ = inputs(20, 0.5)
    d = oracle.DFT(SHAPE, DX, 1.0, model=oracle.WBI, ideal=True, mf=True, kT=1.2, mu=-0.3)
    Hv = d.hess(rho, v)
    h = 1e-5
    fd = (d.eval(rho + h * v)[1] - d.eval(rho - h * v)[1]) / (2 * h)
    assert np.max(np.abs(fd - Hv)) <= 1e-8 * np.max(np.abs(Hv))
    # without the ideal-gas term (full_hessian_ false) the difference is exactly dV v / rho
    H0 = d.hess(rho, v, full=False)
    assert np.max(np.abs(Hv - H0 - d.dV * v / rho)) <= 1e-15 * np.max(np.abs(Hv))
