"""Eigenvalues on the device (include/Eigenvalues.h:16-59, src/Eigenvalues.cpp:46-99): the objective and gradient
of Eigenvalues::calculate_gradients against the oracle to 1e-11 (periodic and fixed background, vshift, change_sign,
full_hessian_ false, squared matrix), and the converged smallest / largest eigenvalue against numpy's eigvalsh of
the dense Hessian.  The oracle's objective is pinned in tests/test_oracle_eigen.py."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def density(shape, seed, rho0=0.4, amp=0.1):
    u = oracle.splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    return rho0 * (1 + amp * (2 * u - 1))


@pytest.mark.parametrize("bnd,vs,sign,squared,ideal,dyn", [(False, False, False, False, True, False),
                                                            (False, True, True, False, True, False),
                                                            (True, True, False, False, True, False),
                                                            (False, False, True, True, False, False),
                                                            (True, False, False, True, True, False),
                                                            (False, True, False, False, True, True),
                                                            (True, False, True, True, True, True)])
def test_objective_matches_oracle(cdft, bnd, vs, sign, squared, ideal, dyn):
    shape, dx = (16, 12, 24), 0.1
    rho = density(shape, 0xE16)
    w = oracle.fmt_weights(shape, dx, 1.0, mode=oracle.WEIGHTS_ACCURATE)
    o = oracle.DFT(shape, dx, 1.0, model=oracle.WBI, mf=True, kT=2.0, accurate_weights=True)
    if bnd:
        o.set_constraints(fixed_background=True)
    s = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mean_field=True, kT=2.0,
                    fixed_background=bnd)
    rng = np.random.default_rng(9)
    x = rng.standard_normal(shape)
    v = rng.standard_normal(shape) if vs else None
    f0, g0 = o.eig_objective(rho, x, v, 1e6, sign, ideal, squared, dyn)
    f1, g1 = s.eigen_objective(rho, x, v, 1e6, sign, ideal, squared, dyn)
    assert abs(f1 - f0) <= 1e-11 * abs(f0)
    assert np.max(np.abs(g1 - g0)) <= 1e-11 * np.max(np.abs(g0))


def test_smallest_and_largest_eigenvalue(cdft):
    shape, dx = (8, 8, 8), 0.2
    rho = density(shape, 0xE17)
    o = oracle.DFT(shape, dx, 1.0, model=oracle.WBI, accurate_weights=True)
    N = rho.size
    H = np.stack([o.hess(rho, np.eye(N)[j].reshape(shape)).ravel() for j in range(N)], axis=1)
    ev = np.linalg.eigvalsh(0.5 * (H + H.T))
    s = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I,
                    weights=oracle.fmt_weights(shape, dx, 1.0, mode=oracle.WEIGHTS_ACCURATE), ideal=True)
    lo = cdft.Eigenvalues(s, rho)
    lo.set_tolerence(1e-13)
    hi = cdft.Eigenvalues(s, rho)
    hi.set_tolerence(1e-13)
    hi.set_change_sign(True)
    e_lo, e_hi = lo.calculate_eigenvector(), hi.calculate_eigenvector()
    print("eigen", e_lo, ev[0], e_hi, ev[-1], lo.num_eval, hi.num_eval)
    assert abs(e_lo - ev[0]) <= 1e-6 * abs(ev[0]) and abs(e_hi - ev[-1]) <= 1e-6 * abs(ev[-1])
    Hv = s.hessian_dot(rho, lo.eigen_vec.reshape(shape)).ravel()
    assert np.linalg.norm(Hv - e_lo * lo.eigen_vec) <= 1e-2 * abs(e_lo)


@pytest.mark.parametrize("dynamic,bnd,shift", [(False, False, 0.0), (False, True, -0.05), (True, False, 0.0)])
def test_arnoldi_largest_eigenpairs(cdft, dynamic, bnd, shift):
    """Arnoldi::determine_largest_eigenpairs (src/Arnoldi.cpp:72-285): the k eigenvalues of A + shift I of largest
    modulus, minus the shift, against numpy's eigenvalues of the dense operator (Hessian, or g H for the DDFT)."""
    shape, dx = (8, 8, 8), 0.2
    rho = density(shape, 0xA70)
    o = oracle.DFT(shape, dx, 1.0, model=oracle.WBI, accurate_weights=True)
    N = rho.size
    mask = np.zeros(shape, bool)
    if bnd:
        mask[0], mask[:, 0], mask[:, :, 0] = True, True, True
    P = np.diag(np.where(mask.ravel(), 0.0, 1.0))
    H = np.stack([o.hess(rho, np.eye(N)[j].reshape(shape)).ravel() for j in range(N)], axis=1)
    A = P @ H @ P
    if dynamic:
        G = np.stack([o.gdotx(rho, np.eye(N)[j].reshape(shape)).ravel() for j in range(N)], axis=1)
        A = P @ G @ A
    act = ~mask.ravel()
    ev = np.linalg.eigvals(A[act][:, act]).real
    ref = ev[np.argsort(-np.abs(ev + shift))][:4]
    s = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I,
                    weights=oracle.fmt_weights(shape, dx, 1.0, mode=oracle.WEIGHTS_ACCURATE), ideal=True,
                    fixed_background=bnd)
    ar = cdft.Arnoldi(s, rho, dynamic=dynamic)
    got = ar.determine_largest_eigenpairs(shift, 4, 16, tol=1e-12)
    print("arnoldi", got, ref)
    assert np.allclose(got, ref, rtol=1e-8, atol=1e-12 * np.abs(ref).max())
    v = ar.get_eigenvector(0)
    assert np.linalg.norm(ar.matrix_dot_v(v, shift) - (got[0] + shift) * v) <= 1e-6 * abs(got[0] + shift)
