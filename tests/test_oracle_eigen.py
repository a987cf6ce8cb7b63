"""Pin of the oracle's Eigenvalues objective (Eigenvalues::calculate_gradients, src/Eigenvalues.cpp:46-70): f is the
Rayleigh quotient of scale * sign * (A + vshift vshift^T / 2), df its gradient 2 (B x - f x) / x.x, with A built
here in numpy from dense oracle Hessian columns (boundary projection, squaring) and checked against central
differences of f."""
import numpy as np
import pytest

import oracle

SHAPE, DX = (6, 6, 6), 0.25


@pytest.mark.parametrize("bnd,vs,sign,squared,full", [(False, False, False, False, True), (False, True, True, False, True),
                                                       (True, True, False, False, True), (False, False, True, True, False)])
def test_objective_is_the_rayleigh_quotient(bnd, vs, sign, squared, full):
    rng = np.random.default_rng(5)
    rho = 0.4 * (1 + 0.1 * (2 * rng.random(SHAPE) - 1))
    o = oracle.DFT(SHAPE, DX, 1.0, model=oracle.WBI, mf=True, kT=1.5)
    op = oracle.DFT(SHAPE, DX, 1.0, model=oracle.WBI, mf=True, kT=1.5)
    if bnd:
        o.set_constraints(fixed_background=True)
    N = rho.size
    H = np.stack([op.hess(rho, np.eye(N)[j].reshape(SHAPE), full=full).ravel() for j in range(N)], axis=1)
    mask = np.zeros(SHAPE, bool)
    if bnd:
        mask[0], mask[:, 0], mask[:, :, 0] = True, True, True
    P = np.diag(np.where(mask.ravel(), 0.0, 1.0))
    A = P @ H @ H @ P if squared else P @ H @ P
    v = rng.standard_normal(N) if vs else None
    scale = 1e3 * (-1 if sign else 1)
    B = scale * (A + (0.5 * np.outer(v, v) if vs else 0.0))
    x = rng.standard_normal(N)
    f, df = o.eig_objective(rho, x.reshape(SHAPE), None if v is None else v.reshape(SHAPE), abs(scale), sign, full,
                            squared)
    x2 = x @ x
    fe = x @ B @ x / x2
    ge = 2 * (B @ x - fe * x) / x2
    assert abs(f - fe) <= 1e-12 * abs(fe)
    assert np.max(np.abs(df.ravel() - ge)) <= 1e-11 * np.max(np.abs(ge))
    if not bnd and not squared:  # central differences along a random direction
        d = rng.standard_normal(N)
        h = 1e-5
        fp, _ = o.eig_objective(rho, (x + h * d).reshape(SHAPE), None if v is None else v.reshape(SHAPE), abs(scale),
                                sign, full, squared)
        fm, _ = o.eig_objective(rho, (x - h * d).reshape(SHAPE), None if v is None else v.reshape(SHAPE), abs(scale),
                                sign, full, squared)
        assert abs((fp - fm) / (2 * h) - df.ravel() @ d) <= 1e-6 * abs(df.ravel() @ d) + 1e-9
