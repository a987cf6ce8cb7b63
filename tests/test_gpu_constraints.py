"""Species constraints on the device against the oracle (Species::beginForceCalculation /
endForceCalculation, src/Species.cpp:64-112; DFT::calculateFreeEnergyAndDerivatives(onlyFex),
src/DFT.cpp:360-425; fixed boundaries of DFT::matrix_dot_v_intern, src/DFT.cpp:441-485, and of
DDFT, src/DDFT.cpp:276-287, 331-337): evaluations to 1e-11, the rescaled density, the fixed-mass
multiplier, FIRE and DDFT trajectories to 1e-10; the oracle's constraints are pinned in
tests/test_oracle_constraints.py."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-11
SHAPE, DX, HSD, KT, MU = (16, 12, 24), 0.1, 1.0, 1.0, 0.7


def density(seed, shape=SHAPE, rho0=0.5, amp=0.1):
    u = oracle.splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    x = [2 * np.pi * g / n for g, n in zip(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), shape)]
    return rho0 * (1 + 0.15 * np.cos(x[0]) * np.sin(x[2]) + amp * (2 * u - 1))


def vext(shape=SHAPE):
    z = np.arange(shape[2]) * DX
    return np.broadcast_to(0.3 * np.exp(-z / 0.3), shape).copy()


def maxrel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def close(a, b, tol):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300)


def pair(cdft, shape=SHAPE, mu=MU, **cons):
    o = oracle.DFT(shape, DX, HSD, model=oracle.WBI, mf=True, kT=KT, mu=mu, vext=vext(shape))
    o.set_constraints(**cons)
    s = cdft.System(shape, DX, HSD, model=cdft.WHITEBEAR_I, weights=oracle.fmt_weights(shape, DX, HSD), ideal=True,
                    mu=mu, vext=vext(shape), mean_field=True, kT=KT, **cons)
    return o, s


def check_eval(o, s, rho):
    p0, d0 = o.eval(rho)
    p1, d1 = s.eval(rho)
    for k in range(5):
        assert abs(p1[k] - p0[k]) <= TOL * abs(p0[0]), (k, p1[k], p0[k])
    assert maxrel(d1, d0) <= TOL
    return d1


@pytest.mark.parametrize("cons", [dict(fixed_mass=1.0), dict(fixed_background=True), dict(homogeneous_boundary=True),
                                  dict(homogeneous_boundary=True, fixed_mass=1.0)],
                         ids=["fixed_mass", "fixed_background", "homogeneous", "homogeneous_fixed_mass"])
def test_constrained_evaluation(cdft, cons):
    rho = density(0x5EED9000)
    if "fixed_mass" in cons:
        cons = dict(cons, fixed_mass=0.9 * rho.sum() * DX ** 3)
    o, s = pair(cdft, **cons)
    check_eval(o, s, rho)
    if "fixed_mass" in cons:
        assert close(s.chem_potential(), o.chem_potential, TOL)
        # device pointers: the density is rescaled in place, like the species density in the reference
        r = cdft.DeviceArray.from_host(rho)
        dF = cdft.DeviceArray(SHAPE)
        s.eval_device(r.ptr, dF.ptr)
        assert maxrel(r.download(), o.density) <= 1e-14
    else:
        assert s.chem_potential() == MU


def test_only_fex(cdft):
    """onlyFex drops the ideal gas and keeps the external field and mu."""
    rho = density(0x5EED9100)
    o = oracle.DFT(SHAPE, DX, HSD, model=oracle.WBI, ideal=False, mf=True, kT=KT, mu=MU, vext=vext())
    _, s = pair(cdft)
    p0, d0 = o.eval(rho)
    p1, d1 = s.eval(rho, only_fex=True)
    assert p1[1] == 0.0
    for k in range(5):
        assert abs(p1[k] - p0[k]) <= TOL * abs(p0[0])
    assert maxrel(d1, d0) <= TOL
    p2, _ = s.eval(rho)  # the flag holds for one call only
    assert p2[1] != 0.0


def test_batch_members_with_fixed_mass(cdft):
    rhos = np.array([density(1), density(2, rho0=0.45)])
    M = 0.9 * rhos[0].sum() * DX ** 3
    w = oracle.fmt_weights(SHAPE, DX, HSD)
    b = cdft.System(SHAPE, DX, HSD, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mus=[MU, MU],
                    vexts=[vext(), vext()], mean_field=True, kT=KT, fixed_mass=M, homogeneous_boundary=True)
    pb, db = b.eval(rhos)
    lam = b.chem_potential()
    for i in range(2):
        o, s = pair(cdft, fixed_mass=M, homogeneous_boundary=True)
        p1, d1 = s.eval(rhos[i])
        assert np.max(np.abs(pb[i] - p1)) <= 1e-13 * abs(p1[0])
        assert maxrel(db[i], d1) <= 1e-13
        assert close(lam[i], s.chem_potential(), 1e-13)


def test_batch_members_with_their_own_fixed_mass(cdft):
    rhos = np.array([density(3), density(4, rho0=0.4)])
    Ms = [0.9 * rhos[0].sum() * DX ** 3, 1.05 * rhos[1].sum() * DX ** 3]
    w = oracle.fmt_weights(SHAPE, DX, HSD)
    b = cdft.System(SHAPE, DX, HSD, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mus=[MU, MU],
                    vexts=[vext(), vext()], mean_field=True, kT=KT, fixed_mass=Ms)
    pb, db = b.eval(rhos)
    lam = b.chem_potential()
    for i in range(2):
        o, _ = pair(cdft, fixed_mass=Ms[i])
        p0, d0 = o.eval(rhos[i])
        assert np.max(np.abs(pb[i] - p0)) <= TOL * abs(p0[0])
        assert maxrel(db[i], d0) <= TOL
        assert close(lam[i], o.chem_potential, TOL)


def test_fire_with_fixed_mass_replays_the_oracle(cdft):
    shape = (16, 16, 16)
    rho0 = density(11, shape)
    M = 0.95 * rho0.sum() * DX ** 3
    o, s = pair(cdft, shape=shape, mu=0.0, fixed_mass=M)
    fo, fg = oracle.Fire(o, rho0), cdft.Fire(s, rho0)
    for i in range(30):
        rc0, a = fo.step()
        rc1, b = fg.step()
        assert rc0 == rc1 == 0, (i, rc0, rc1)
        assert close(a["F"], b["F"], 1e-10), (i, a["F"], b["F"])
        assert close(a["dt"], b["dt"], 1e-12) and close(a["alpha"], b["alpha"], 1e-12)
        assert close(a["fmax_raw"], b["fmax_raw"], 1e-9)
    assert close(s.chem_potential(), o.chem_potential, 1e-9)
    r0, r1 = fo.state()[0], fg.state()[0]
    assert np.abs(r0 - r1).max() <= 1e-10 * np.abs(r0).max()
    assert abs(r1.sum() * DX ** 3 - M) <= 1e-12 * M


@pytest.mark.parametrize("cons", [dict(fixed_background=True), dict(fixed_mass=1.0)], ids=["fixed_background", "fixed_mass"])
def test_ddft_with_constraints_replays_the_oracle(cdft, cons):
    shape = (16, 16, 16)
    rho0 = density(21, shape)
    if "fixed_mass" in cons:
        cons = dict(fixed_mass=rho0.sum() * DX ** 3)
    o, s = pair(cdft, shape=shape, mu=0.0, **cons)
    do, dg = oracle.DDFT(o, rho0), cdft.DDFT(s, rho0)
    for i in range(3):
        a, b = do.step(), dg.step()
        assert a["calls"] == b["calls"] and a["restarts"] == b["restarts"]
        assert close(a["F"], b["F"], 1e-10) and close(a["dt"], b["dt"], 1e-12)
    r0, r1 = do.density(), dg.density()
    assert np.abs(r0 - r1).max() <= 1e-10 * np.abs(r0).max()
    if "fixed_background" in cons:
        bnd = np.zeros(shape, dtype=bool)
        bnd[0], bnd[:, 0], bnd[:, :, 0] = True, True, True
        assert np.array_equal(r1[bnd], rho0[bnd])


def test_hessian_with_fixed_background(cdft):
    rho = density(31)
    o, s = pair(cdft, fixed_background=True)
    v = 2 * oracle.splitmix64_uniform(int(np.prod(SHAPE)), 33).reshape(SHAPE) - 1
    with pytest.raises(cdft.CdftError):
        s.hessian_dot(rho, v)
    v[0], v[:, 0], v[:, :, 0] = 0.0, 0.0, 0.0
    h0, h1 = o.hess(rho, v), s.hessian_dot(rho, v)
    assert maxrel(h1, h0) <= TOL
    assert np.all(h1[0] == 0) and np.all(h1[:, 0] == 0) and np.all(h1[:, :, 0] == 0)


def test_fixed_background_with_fixed_mass_is_rejected(cdft):
    with pytest.raises(cdft.CdftError):
        cdft.System(SHAPE, DX, HSD, ideal=True, fixed_mass=1.0, fixed_background=True)
