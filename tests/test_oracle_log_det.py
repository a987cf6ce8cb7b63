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


def _dense(f, N):
    return np.stack([f(np.eye(N)[j].reshape(SHAPE)).ravel() for j in range(N)], axis=1)


def test_mobility_operator():
    """g . x of DDFT::g_dot_x (src/DDFT.cpp:320-351): symmetric, negative semi-definite, constants in its null space
    and its columns summing to zero (mass conservation), rho0 times the discrete Laplacian at uniform density."""
    o, rho = _dft()
    N = rho.size
    G = _dense(lambda x: o.gdotx(rho, x), N)
    assert np.max(np.abs(G - G.T)) <= 1e-12 * np.max(np.abs(G))
    assert np.max(np.linalg.eigvalsh(G)) <= 1e-10 * np.max(np.abs(G))
    assert np.max(np.abs(G @ np.ones(N))) <= 1e-12 * np.max(np.abs(G)) and np.max(np.abs(G.sum(0))) <= 1e-10
    rho0 = np.full(SHAPE, 0.4)
    k = 2 * np.pi * 2 / (SHAPE[1] * DX)
    y = np.arange(SHAPE[1]) * DX
    wave = np.broadcast_to(np.cos(k * y)[None, :, None], SHAPE).copy()
    lam = 0.4 * 2 * (np.cos(k * DX) - 1) / DX ** 2
    assert np.max(np.abs(o.gdotx(rho0, wave) - lam * wave)) <= 1e-12 * abs(lam)


def test_mobility_operator_central_differences():
    """g . x with central differences (DDFT(dft, showGraphics, true), src/DDFT.cpp:325, 345): symmetric, negative
    semi-definite, mass conserving; at uniform density rho0 (x_{i+2} - 2 x_i + x_{i-2}) / (4 dx^2) per axis, whose
    Fourier eigenvalue is rho0 (cos(2 k dx) - 1) / (2 dx^2); and different from the forward-difference operator."""
    o, rho = _dft()
    o.set_central_differences(True)
    N = rho.size
    G = _dense(lambda x: o.gdotx(rho, x), N)
    assert np.max(np.abs(G - G.T)) <= 1e-12 * np.max(np.abs(G))
    assert np.max(np.linalg.eigvalsh(G)) <= 1e-10 * np.max(np.abs(G))
    assert np.max(np.abs(G @ np.ones(N))) <= 1e-12 * np.max(np.abs(G)) and np.max(np.abs(G.sum(0))) <= 1e-10
    rho0 = np.full(SHAPE, 0.4)
    k = 2 * np.pi * 1 / (SHAPE[1] * DX)
    y = np.arange(SHAPE[1]) * DX
    wave = np.broadcast_to(np.cos(k * y)[None, :, None], SHAPE).copy()
    lam = 0.4 * (np.cos(2 * k * DX) - 1) / (2 * DX ** 2)
    assert np.max(np.abs(o.gdotx(rho0, wave) - lam * wave)) <= 1e-12 * abs(lam)
    o.set_central_differences(False)
    G1 = _dense(lambda x: o.gdotx(rho, x), N)
    assert np.max(np.abs(G1 - G)) > 0.1 * np.max(np.abs(G))


@pytest.mark.parametrize("bnd,squared", [(False, False), (True, False), (False, True)])
def test_dynamic_log_det_is_the_dense_estimator(bnd, squared):
    """Log_Det on the DDFT dynamical matrix (DDFT::matrix_dot_v_intern, src/DDFT.cpp:353-361; is_dynamic negation,
    src/Dynamical_Matrix.cpp:48 and src/Log_Det.cpp:124, 139), zero mode shifted: the estimator evaluated with the
    dense, non-symmetric operator -P G P H P (squared: -(P G P H P)^2) by the matrix three-term recurrence."""
    o, rho = _dft(bnd)
    op, _ = _dft(False)
    N = rho.size
    mask = _boundary_mask() if bnd else np.zeros(N, bool)
    P = np.diag(np.where(mask, 0.0, 1.0))
    H = _dense(lambda x: op.hess(rho, x), N)
    G = _dense(lambda x: op.gdotx(rho, x), N)
    M = P @ G @ P @ H @ P
    A = -(M @ M) if squared else M
    assert np.max(np.abs(_dense(lambda x: o.dynamic_dot(rho, x, squared=squared), N) - A)) <= 1e-10 * np.max(np.abs(A))
    B = -A  # Log_Det negates the products of a dynamical matrix
    ev = np.sort(np.linalg.eigvals(B[~mask][:, ~mask]).real)
    lam_min, lam_max = 0.5 * ev[ev > 1e-9 * ev.max()].min(), 1.2 * ev.max()
    lam_mid = 0.5 * (lam_min + lam_max)
    B = B + lam_mid / N * np.ones((N, N))
    order, seed, ns = 16, 5, 3
    sc = 1.0 / (lam_min + lam_max)
    a, b = lam_min * sc, lam_max * sc
    X = (2.0 / (b - a)) * sc * B - (b + a) / (b - a) * np.eye(N)
    c = oracle.cheb_log_coeffs(order, lam_max, lam_min)
    T0, T1 = np.eye(N), X
    pX = c[0] * T0 + c[1] * T1
    for k in range(2, order + 1):
        T0, T1 = T1, 2 * X @ T1 - T0
        pX += c[k] * T1
    V = oracle.rademacher(seed, N, ns) * np.where(mask, 0.0, 1.0)
    q = np.einsum("si,ij,sj->s", V, pX, V)
    expect = q.mean() - (N - mask.sum()) * np.log(sc) - np.log(lam_mid)
    ld, _ = o.log_det(rho, order, lam_max, lam_min, seed, ns, has_zero=True, squared=squared, dynamic=True)
    assert abs(ld - expect) <= 1e-9 * abs(expect)
