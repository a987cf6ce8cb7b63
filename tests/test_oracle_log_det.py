"""Pins of the oracle's stochastic log-determinant (Log_Det, src/Log_Det.cpp:22-166).

* probes: numpy's legacy MT19937 reproduces the standard's known answer (10000th output of the default-seeded
  mt19937 is 4123659995), and the oracle's +-1 probes are the top bits of those outputs (libstdc++'s
  uniform_int_distribution<>(0, 1) on a 32-bit engine);
* coefficients: numpy's Chebyshev interpolant at the same first-kind nodes;
* estimator: the mean of v^T p(X) v over the same probes, with p(X) evaluated on the dense operator (Hessian
  columns from the oracle, boundary projection, rank-one zero-mode shift and squaring built here in numpy);
* sanity: the estimate agrees with numpy's slogdet within the sampling error.
"""
import numpy as np
import pytest

import oracle

SHAPE, DX = (6, 6, 6), 0.25


def test_probes_are_mt19937_top_bits():
    mt = np.random.RandomState(5489).randint(0, 2 ** 32, size=10000, dtype=np.uint32)
    assert int(mt[-1]) == 4123659995
    for seed in (1, 42, 123456789):
        raw = np.random.RandomState(seed).randint(0, 2 ** 32, size=3 * 500, dtype=np.uint32)
        expect = np.where(raw >> 31, 1.0, -1.0).reshape(3, 500)
        assert np.array_equal(oracle.rademacher(seed, 500, 3), expect)


@pytest.mark.parametrize("order,lmax,lmin", [(5, 2.0, 0.1), (20, 0.05, 1e-3), (60, 10.0, 0.5)])
def test_chebyshev_coefficients_are_the_interpolant(order, lmax, lmin):
    c = oracle.cheb_log_coeffs(order, lmax, lmin)
    sc = 1.0 / (lmin + lmax)
    a, b = lmin * sc, lmax * sc
    ref = np.polynomial.chebyshev.chebinterpolate(lambda x: np.log(0.5 * (b - a) * x + 0.5 * (b + a)), order)
    assert np.max(np.abs(c - ref)) <= 1e-13 * np.max(np.abs(ref))


def _dft(bnd=False):
    o = oracle.DFT(SHAPE, DX, 1.0, model=oracle.WBI, mf=True, kT=1.5, mu=0.0)
    if bnd:
        o.set_constraints(fixed_background=True)
    rng = np.random.default_rng(3)
    rho = 0.4 * (1 + 0.1 * (2 * rng.random(SHAPE) - 1))
    return o, rho


def _boundary_mask():
    m = np.zeros(SHAPE, bool)
    m[0, :, :] = m[:, 0, :] = m[:, :, 0] = True
    return m.ravel()


def _dense_hessian(o, rho, full):
    N = rho.size
    H = np.zeros((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        H[:, j] = o.hess(rho, e.reshape(SHAPE), full=full).ravel()
    return H


@pytest.mark.parametrize("bnd,has_zero,full,squared", [(False, False, True, False), (False, True, True, False),
                                                        (False, False, False, True), (True, True, True, False)])
def test_log_det_is_the_dense_estimator(bnd, has_zero, full, squared):
    o, rho = _dft(bnd)
    N = rho.size
    mask = _boundary_mask() if bnd else np.zeros(N, bool)
    P = np.diag(np.where(mask, 0.0, 1.0))
    H = _dense_hessian(o, rho, full) if not bnd else None
    if bnd:  # columns of boundary sites are never reached: the input is projected first
        o2, _ = _dft(False)
        H = P @ _dense_hessian(o2, rho, full) @ P
    A = P @ H @ H @ P if squared else P @ H @ P
    ev = np.linalg.eigvalsh(A[~mask][:, ~mask])
    lam_min, lam_max = 0.8 * ev.min(), 1.2 * ev.max()
    lam_mid = 0.5 * (lam_min + lam_max)
    if has_zero:
        A = A + lam_mid / N * np.ones((N, N))
    order, seed, ns = 24, 77, 4
    sc = 1.0 / (lam_min + lam_max)
    a, b = lam_min * sc, lam_max * sc
    X = (2.0 / (b - a)) * sc * A - (b + a) / (b - a) * np.eye(N)
    c = oracle.cheb_log_coeffs(order, lam_max, lam_min)
    w, Q = np.linalg.eigh(0.5 * (X + X.T))
    pX = (Q * np.polynomial.chebyshev.chebval(w, c)) @ Q.T
    V = oracle.rademacher(seed, N, ns) * np.where(mask, 0.0, 1.0)
    q = np.einsum("si,ij,sj->s", V, pX, V)
    expect = q.mean() - (N - mask.sum()) * np.log(sc) - (np.log(lam_mid) if has_zero else 0.0)
    var = np.sqrt(abs(np.mean(q ** 2) - q.mean() ** 2))
    ld, v = o.log_det(rho, order, lam_max, lam_min, seed, ns, has_zero=has_zero, full=full, squared=squared)
    assert abs(ld - expect) <= 1e-10 * abs(expect)
    assert abs(v - var) <= 1e-8 * max(var, 1.0)


def test_log_det_estimates_slogdet():
    o, rho = _dft()
    H = _dense_hessian(o, rho, True)
    ev = np.linalg.eigvalsh(0.5 * (H + H.T))
    sign, exact = np.linalg.slogdet(H)
    assert sign > 0
    ns = 200
    ld, var = o.log_det(rho, 40, 1.2 * ev.max(), 0.8 * ev.min(), 11, ns)
    assert abs(ld - exact) <= 4 * var / np.sqrt(ns) + 1e-8 * abs(exact)
