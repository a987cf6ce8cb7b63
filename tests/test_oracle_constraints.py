"""Pins the oracle's species constraints (Species::beginForceCalculation / endForceCalculation,
src/Species.cpp:64-112; fixed boundaries in DFT::matrix_dot_v_intern, src/DFT.cpp:441-485, and in
DDFT, src/DDFT.cpp:276-287, 331-337) before they check the device.

Fixed mass: the density is rescaled to the mass, the constrained force is orthogonal to it, and on
zero-sum directions it is still the derivative of F (central differences), so the projection removes
exactly a constant times dV; for a uniform density the multiplier is the bulk chemical potential.
Fixed background / homogeneous boundary: the boundary planes ix, iy or iz = 0 (their count
NxNy + NxNz + NyNz - Nx - Ny - Nz + 1 on an anisotropic grid) are zeroed / replaced by their mean,
everything else untouched."""
import numpy as np
import pytest

import oracle

SHAPE, DX, KT = (12, 10, 14), 0.1, 1.0


def density(seed, rho0=0.5):
    rng = np.random.default_rng(seed)
    x = [2 * np.pi * g / n for g, n in zip(np.meshgrid(*[np.arange(n) for n in SHAPE], indexing="ij"), SHAPE)]
    return rho0 * (1 + 0.2 * np.cos(x[0]) * np.sin(x[2]) + 0.1 * np.cos(x[1]) + 0.05 * (2 * rng.random(SHAPE) - 1))


def vext():
    z = np.arange(SHAPE[2]) * DX
    return np.broadcast_to(0.3 * np.exp(-z / 0.3), SHAPE).copy()


def make(**kw):
    o = oracle.DFT(SHAPE, DX, 1.0, model=oracle.WBI, mf=True, kT=KT, mu=0.7, vext=vext())
    o.set_constraints(**kw)
    return o


def boundary_mask(shape=SHAPE):
    m = np.zeros(shape, dtype=bool)
    m[0, :, :] = m[:, 0, :] = m[:, :, 0] = True
    return m


def test_boundary_count():
    Nx, Ny, Nz = SHAPE
    assert boundary_mask().sum() == Nx * Ny + Nx * Nz + Ny * Nz - Nx - Ny - Nz + 1


def test_fixed_mass_rescales_and_projects():
    rho = density(1)
    dV = DX ** 3
    M = 0.9 * rho.sum() * dV
    o = make(fixed_mass=M)
    parts, dF = o.eval(rho)
    r = o.density
    assert abs(r.sum() * dV - M) <= 1e-13 * M
    assert np.max(np.abs(r - rho * (M / (rho.sum() * dV)))) <= 1e-14 * r.max()
    # mu is zero in the external term: F_ext = sum V rho dV
    assert abs(parts[4] - np.sum(vext() * r) * dV) <= 1e-13 * abs(parts[4])
    # the constrained force is orthogonal to the density ...
    assert abs(np.sum(dF * r)) <= 1e-12 * np.sum(np.abs(dF * r))
    # ... and still the derivative of F along zero-sum directions (which keep the mass)
    rng = np.random.default_rng(2)
    for _ in range(2):
        v = 2 * rng.random(SHAPE) - 1
        v -= v.mean()
        h = 1e-6
        fd = (make(fixed_mass=M).eval(r + h * v)[0][0] - make(fixed_mass=M).eval(r - h * v)[0][0]) / (2 * h)
        assert abs(fd - np.sum(dF * v)) <= 1e-7 * np.sum(np.abs(dF * v)), (fd, np.sum(dF * v))
    # the multiplier is the constant the projection removed: dF_free = dF + mu dV
    free = oracle.DFT(SHAPE, DX, 1.0, model=oracle.WBI, mf=True, kT=KT, mu=0.0, vext=vext())
    _, dF0 = free.eval(r)
    assert np.max(np.abs(dF0 - dF - o.chem_potential * dV)) <= 1e-12 * np.max(np.abs(dF0))


def test_fixed_mass_multiplier_is_bulk_mu():
    rb, dV = 0.6, DX ** 3
    o = oracle.DFT(SHAPE, DX, 1.0, model=oracle.WBI, mf=True, kT=KT, mu=0.7, accurate_weights=True)
    o.set_constraints(fixed_mass=0.5 * np.prod(SHAPE) * dV)
    _, dF = o.eval(np.full(SHAPE, rb))
    assert np.max(np.abs(dF)) <= 1e-12
    _, muex = oracle.bulk(oracle.WBI, 0.5, 1.0)
    assert abs(o.chem_potential - (np.log(0.5) + muex + o.a_vdw * 0.5)) <= 1e-12


def test_fixed_background_and_homogeneous_boundary():
    rho = density(3)
    _, dF0 = make().eval(rho)
    b = boundary_mask()
    _, dFb = make(fixed_background=True).eval(rho)
    assert np.all(dFb[b] == 0.0) and np.array_equal(dFb[~b], dF0[~b])
    _, dFh = make(homogeneous_boundary=True).eval(rho)
    assert np.array_equal(dFh[~b], dF0[~b])
    assert np.max(np.abs(dFh[b] - dF0[b].mean())) <= 1e-15 * np.abs(dF0).max()
    # homogeneous boundary with a fixed mass: constant on the boundary, orthogonal to the density
    o = make(homogeneous_boundary=True, fixed_mass=0.8 * rho.sum() * DX ** 3)
    _, dFm = o.eval(rho)
    assert np.ptp(dFm[b]) <= 1e-15 * np.abs(dFm).max()
    assert abs(np.sum(dFm * o.density)) <= 1e-12 * np.sum(np.abs(dFm * o.density))
    with pytest.raises(ValueError):
        make(fixed_background=True, fixed_mass=1.0)


def test_hessian_with_fixed_background():
    rho = density(4)
    b = boundary_mask()
    v = 2 * np.random.default_rng(5).random(SHAPE) - 1
    with pytest.raises(ValueError):
        make(fixed_background=True).hess(rho, v)
    v[b] = 0.0
    h0 = make().hess(rho, v)
    h1 = make(fixed_background=True).hess(rho, v)
    assert np.all(h1[b] == 0.0) and np.array_equal(h1[~b], h0[~b])


def test_fire_keeps_the_mass():
    rho = density(6)
    M = rho.sum() * DX ** 3
    o = make(fixed_mass=M)
    f = oracle.Fire(o, rho)
    F0 = o.eval(rho)[0][0]
    for _ in range(30):
        rc, info = f.step()
        assert rc == 0
        r = f.state()[0]
        assert abs(r.sum() * DX ** 3 - M) <= 1e-12 * M
    assert info["F"] < F0


def test_ddft_fixed_background_keeps_the_boundary():
    rho = density(7)
    o = make(fixed_background=True)
    d = oracle.DDFT(o, rho)
    for _ in range(3):
        d.step()
    r = d.density()
    b = boundary_mask()
    assert np.array_equal(r[b], rho[b])
    assert np.max(np.abs(r[~b] - rho[~b])) > 1e-6


def test_fire_fixed_direction_reflects_the_force():
    """Minimizer::getDF_DX with a fixed direction e (src/Minimizer.cpp:103-110): dF' = dF - 2 (e.dF) e is the
    reflection of the force in the plane orthogonal to e (|e| = 1): the e component changes sign, the rest stays."""
    rho = density(8)
    o = make()
    e = np.random.default_rng(9).standard_normal(SHAPE)
    e /= np.linalg.norm(e)
    f = oracle.Fire(o, rho)
    f.set_fixed_direction(e)
    f.reset(rho)
    r, _, v, d1 = f.state()
    assert np.all(v == 0.0)
    _, d0 = make().eval(r)
    assert abs(np.sum(e * d1) + np.sum(e * d0)) <= 1e-12 * np.linalg.norm(d0)
    perp0 = d0 - np.sum(e * d0) * e
    perp1 = d1 - np.sum(e * d1) * e
    assert np.max(np.abs(perp1 - perp0)) <= 1e-12 * np.abs(d0).max()
    assert abs(np.linalg.norm(d1) - np.linalg.norm(d0)) <= 1e-12 * np.linalg.norm(d0)


def test_fire_reset_keeps_dt_and_restarts_from_the_density():
    rho = density(10)
    o = make()
    f = oracle.Fire(o, rho)
    for _ in range(8):
        _, info = f.step()
    r = f.state()[0]
    assert f.reset(r) == 0
    r2, _, v, _ = f.state()
    assert np.all(v == 0.0) and np.max(np.abs(r2 - r)) <= 1e-12 * r.max()
    _, info2 = f.step()
    assert info2["dt"] >= info["dt"] * 0.5  # dt is kept (a fresh minimizer would restart at 1e-3)
