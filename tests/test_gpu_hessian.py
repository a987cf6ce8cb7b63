"""Hessian-vector product on the device (dual-number second derivatives of Phi)
against the oracle's restatement of FMT::add_second_derivative /
DFT::matrix_dot_v_intern (src/FMT.cpp:181-280, 437-567; src/DFT.cpp:435-488),
which tests/test_oracle_hessian.py pins by central differences.

Two independent evaluations: the oracle uses the reference's explicit
second-derivative formulas per model; the device differentiates its own
first-derivative code in forward mode.  Same inputs, same weights.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-11
SHAPE, DX = (16, 12, 24), 0.1


def maxrel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def inputs(seed, shape=SHAPE, rho0=0.6):
    u = oracle.splitmix64_uniform(2 * int(np.prod(shape)), seed)
    n = int(np.prod(shape))
    x = np.meshgrid(*[np.arange(k) for k in shape], indexing="ij")
    wave = np.cos(2 * np.pi * x[0] / shape[0]) * np.sin(2 * np.pi * x[2] / shape[2])
    rho = rho0 * (1 + 0.25 * wave + 0.05 * (2 * u[:n].reshape(shape) - 1))
    return rho, 2 * u[n:].reshape(shape) - 1


MODELS = [(0, "ROSENFELD"), (1, "RSLT"), (2, "ESFMT"), (3, "WHITEBEAR_I"), (4, "WHITEBEAR_II")]


@pytest.mark.parametrize("model,name", MODELS)
def test_fmt_hessian_matches_oracle(cdft, model, name):
    rho, v = inputs(0x5EED7000 + model)
    w = oracle.fmt_weights(SHAPE, DX, 1.0)
    A, B = (1.0, -0.5) if model == 2 else (1.0, 0.0)
    o = oracle.DFT(SHAPE, DX, 1.0, model=model, A=A, B=B, ideal=False)
    ref = o.hess(rho, v)
    s = cdft.System(SHAPE, DX, 1.0, model=getattr(cdft, name), A=A, B=B, weights=w, ideal=False)
    got = s.fmt_hessian_dot(rho, v)
    assert maxrel(got, ref) <= TOL, maxrel(got, ref)
    # accumulate: adds to what out holds (Species d2F accumulation)
    base = np.full(SHAPE, 0.25)
    acc = s.fmt_hessian_dot(rho, v, out=base.copy(), accumulate=True)
    assert maxrel(acc - base, ref) <= TOL


def test_dft_hessian_matches_oracle(cdft):
    rho, v = inputs(0x5EED7100, rho0=0.5)
    kT, mu = 0.8, -1.3
    z = np.arange(SHAPE[2]) * DX
    vext = np.broadcast_to(np.exp(-z / 0.2), SHAPE).copy()
    w = oracle.fmt_weights(SHAPE, DX, 1.0)
    o = oracle.DFT(SHAPE, DX, 1.0, model=oracle.WBI, mf=True, kT=kT, mu=mu, vext=vext)
    s = cdft.System(SHAPE, DX, 1.0, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mu=mu, vext=vext,
                    mean_field=True, kT=kT)
    for full in (True, False):
        ref = o.hess(rho, v, full=full)
        got = s.hessian_dot(rho, v, ideal=full)
        assert maxrel(got, ref) <= TOL, (full, maxrel(got, ref))


def test_hessian_symmetric_and_linear_at_256_axis(cdft):
    """Size-independent properties on a long axis: u.(Hv) = v.(Hu), H(a u + b v) = a Hu + b Hv, and
    the directional derivative of dF."""
    shape = (32, 24, 256)
    rho, v = inputs(0x5EED7200, shape)
    u = np.roll(v, 5, axis=2) * np.cos(np.arange(shape[0]))[:, None, None]
    s = cdft.System(shape, 0.05, 1.0, model=cdft.WHITEBEAR_I, ideal=True)
    Hv = s.hessian_dot(rho, v)
    Hu = s.hessian_dot(rho, u)
    assert abs(np.sum(u * Hv) - np.sum(v * Hu)) <= 1e-12 * np.sum(np.abs(u * Hv))
    Hc = s.hessian_dot(rho, 2.0 * u - 0.5 * v)
    assert maxrel(Hc, 2.0 * Hu - 0.5 * Hv) <= 1e-13
    h = 1e-5
    fd = (s.eval(rho + h * v)[1] - s.eval(rho - h * v)[1]) / (2 * h)
    assert maxrel(fd, Hv) <= 1e-7


def test_batched_hessian_matches_members(cdft):
    nb = 3
    pairs = [inputs(0x5EED7300 + b) for b in range(nb)]
    rhos = np.array([p[0] for p in pairs])
    vs = np.array([p[1] for p in pairs])
    bs = cdft.System(SHAPE, DX, 1.0, mus=[0.1, 0.2, 0.3], model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=1.0)
    H = bs.hessian_dot(rhos, vs)
    for b in range(nb):
        s1 = cdft.System(SHAPE, DX, 1.0, mu=0.1 * (b + 1), model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=1.0)
        assert maxrel(H[b], s1.hessian_dot(rhos[b], vs[b])) <= 1e-14


def test_hessian_needs_fp64(cdft):
    s = cdft.System(SHAPE, DX, 1.0, precision=cdft.MIXED)
    rho, v = inputs(1)
    with pytest.raises(cdft.CdftError):
        s.hessian_dot(rho, v)
