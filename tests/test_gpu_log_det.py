"""Stochastic log-determinant of the Hessian on the device (Log_Det, src/Log_Det.cpp:22-166) against the oracle:
the same probes (std::mt19937 + uniform_int_distribution<>(0, 1)), log det and variance to 1e-10, periodic and
fixed-background species, the zero-mode shift, full_hessian_ false and use_squared_matrix_.  The oracle's
estimator is pinned in tests/test_oracle_log_det.py."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def density(shape, seed, rho0=0.5, amp=0.1):
    u = oracle.splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    return rho0 * (1 + amp * (2 * u - 1))


def rel(a, b):
    return abs(a - b) / abs(b)


def variance_close(v1, v0, ld, n_active, lmax, lmin, has_zero, tol=1e-10):
    """The variance is sqrt|<q^2> - <q>^2| over the per-probe values q = v.u: its square carries the rounding of
    <q^2> ~ <q>^2, so it is compared as |v1^2 - v0^2| <= tol <q>^2 (<q> recovered from the log det)."""
    q = ld + n_active * np.log(1.0 / (lmin + lmax)) + (np.log(0.5 * (lmax + lmin)) if has_zero else 0.0)
    return abs(v1 * v1 - v0 * v0) <= tol * q * q


CASES = [
    # shape, mf/kT, fixed background, has_zero, ideal, squared, order, samples, (lam_max, lam_min)
    ((16, 16, 16), None, False, False, True, False, 20, 3, (0.05, 5e-4)),
    ((16, 12, 24), 3.0, False, True, True, False, 16, 2, (0.05, 5e-4)),
    ((16, 16, 16), 3.0, True, True, True, False, 12, 2, (0.05, 5e-4)),
    ((12, 16, 16), None, False, False, False, True, 10, 2, (2.5e-3, 1e-8)),
    ((16, 16, 16), None, True, False, True, True, 10, 2, (2.5e-3, 2.5e-7)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_log_det_matches_oracle(cdft, case):
    shape, kT, bnd, has_zero, ideal, squared, order, ns, (lmax, lmin) = CASES[case]
    dx = 0.1
    rho = density(shape, 0x1D0 + case)
    o = oracle.DFT(shape, dx, 1.0, model=oracle.WBI, mf=kT is not None, kT=kT or 1.0, accurate_weights=True)
    if bnd:
        o.set_constraints(fixed_background=True)
    ld0, var0 = o.log_det(rho, order, lmax, lmin, 1234 + case, ns, has_zero=has_zero, full=ideal, squared=squared)
    s = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I, weights=oracle.fmt_weights(shape, dx, 1.0,
                    mode=oracle.WEIGHTS_ACCURATE), ideal=True, mean_field=kT is not None, kT=kT or 1.0,
                    fixed_background=bnd)
    ld1, var1 = s.log_det(rho, order, lmax, lmin, 1234 + case, ns, has_zero_eigenvalue=has_zero, ideal=ideal,
                          squared=squared)
    print("case", case, "oracle", ld0, var0, "gpu", ld1, var1)
    assert rel(ld1, ld0) <= 1e-10
    n_active = rho.size - (shape[0] * shape[1] + shape[0] * shape[2] + shape[1] * shape[2] - sum(shape) + 1 if bnd else 0)
    assert variance_close(var1, var0, ld0, n_active, lmax, lmin, has_zero)


def test_log_det_64_matches_oracle(cdft):
    shape, dx = (64, 64, 64), 0.05
    rho = density(shape, 0x1DF, rho0=0.4, amp=0.05)
    o = oracle.DFT(shape, dx, 1.0, model=oracle.WBI, accurate_weights=True)
    ld0, var0 = o.log_det(rho, 6, 0.01, 5e-5, 99, 1)
    s = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I,
                    weights=oracle.fmt_weights(shape, dx, 1.0, mode=oracle.WEIGHTS_ACCURATE), ideal=True)
    ld1, var1 = s.log_det(rho, 6, 0.01, 5e-5, 99, 1)
    assert rel(ld1, ld0) <= 1e-10 and var1 == 0.0 and var0 == 0.0


def test_log_det_rejects_bad_arguments(cdft):
    s = cdft.System((16, 16, 16), 0.1, 1.0, model=cdft.WHITEBEAR_I, ideal=True)
    rho = density((16, 16, 16), 5)
    for args in ((0, 0.05, 5e-4), (10, 5e-4, 0.05), (10, 0.05, -1.0)):
        with pytest.raises(cdft.CdftError):
            s.log_det(rho, *args, 1, 1)


def test_dynamic_matrix_central_differences(cdft):
    """The dynamical matrix of a DDFT built with central differences (CDFT_CENTRAL_DIFFERENCES) vs the oracle."""
    shape, dx = (16, 12, 24), 0.1
    rho = density(shape, 0x1E1)
    w = oracle.fmt_weights(shape, dx, 1.0, mode=oracle.WEIGHTS_ACCURATE)
    o = oracle.DFT(shape, dx, 1.0, model=oracle.WBI, mf=True, kT=3.0, accurate_weights=True)
    o.set_central_differences(True)
    s = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mean_field=True, kT=3.0)
    v = np.random.default_rng(5).standard_normal(shape)
    for squared in (False, True):
        y0 = o.dynamic_dot(rho, v, squared=squared)
        y1 = s.dynamic_dot(rho, v, squared=squared, central=True)
        assert np.max(np.abs(y1 - y0)) <= 1e-11 * np.max(np.abs(y0))


@pytest.mark.parametrize("bnd,squared", [(False, False), (True, False), (False, True)])
def test_dynamic_matrix_matches_oracle(cdft, bnd, squared):
    """DDFT::matrix_dot_v_intern (src/DDFT.cpp:353-361) = g . (H v) on the device vs the oracle (pinned in
    tests/test_oracle_log_det.py), and the log det of the dynamical matrix (products negated, zero mode shifted)."""
    shape, dx = (16, 12, 24), 0.1
    rho = density(shape, 0x1E0)
    w = oracle.fmt_weights(shape, dx, 1.0, mode=oracle.WEIGHTS_ACCURATE)
    o = oracle.DFT(shape, dx, 1.0, model=oracle.WBI, mf=True, kT=3.0, accurate_weights=True)
    if bnd:
        o.set_constraints(fixed_background=True)
    s = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mean_field=True, kT=3.0,
                    fixed_background=bnd)
    v = np.random.default_rng(4).standard_normal(shape)
    y0 = o.dynamic_dot(rho, v, squared=squared)
    y1 = s.dynamic_dot(rho, v, squared=squared)
    assert np.max(np.abs(y1 - y0)) <= 1e-11 * np.max(np.abs(y0))
    lmax, lmin = (1600.0, 1e-4) if squared else (40.0, 5e-3)
    ld0, _ = o.log_det(rho, 10, lmax, lmin, 21, 2, has_zero=True, squared=squared, dynamic=True)
    ld1, _ = s.log_det(rho, 10, lmax, lmin, 21, 2, has_zero_eigenvalue=True, squared=squared, dynamic=True)
    assert rel(ld1, ld0) <= 1e-10
