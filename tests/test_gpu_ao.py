"""FMT with an Asakura-Oosawa polymer on the device against the oracle's restatement of
FMT_AO_Species (include/Species.h:380-461, src/FMT_Species.cpp:647-738, src/FMT.cpp:393-422),
whose force tests/test_oracle_ao.py pins as the derivative of its free energy."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-11
SHAPE, DX, HSD, RP, RHO_P = (16, 12, 24), 0.1, 1.0, 0.3, 0.4
MODELS = [(0, "ROSENFELD"), (1, "RSLT"), (2, "ESFMT"), (3, "WHITEBEAR_I"), (4, "WHITEBEAR_II")]


def maxrel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def density(seed, shape=SHAPE, rho0=0.45):
    u = oracle.splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    x = [2 * np.pi * g / n for g, n in zip(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), shape)]
    return rho0 * (1 + 0.25 * np.cos(x[0]) * np.cos(x[1]) + 0.2 * np.sin(x[1] + x[2]) + 0.05 * (2 * u - 1))


@pytest.mark.parametrize("model,name", MODELS)
def test_fmt_with_ao_matches_oracle(cdft, model, name):
    rho = density(0x5EED8000 + model)
    w = oracle.fmt_weights(SHAPE, DX, HSD)
    wp = oracle.fmt_weights(SHAPE, DX, 2 * RP)
    A, B = (1.0, -0.5) if model == 2 else (1.0, 0.0)
    o = oracle.FMT(model, SHAPE, DX, HSD, weights=w, A=A, B=B)
    o.set_ao(RP, RHO_P, ao_weights=wp)
    F0, d0 = o.eval(rho)
    s = cdft.System(SHAPE, DX, HSD, model=getattr(cdft, name), A=A, B=B, weights=w, ao_Rp=RP, ao_reservoir=RHO_P,
                    ao_weights=wp)
    F1, d1 = s.fmt_eval(rho)
    assert abs(F1 - F0) <= TOL * abs(F0), (F1, F0)
    assert maxrel(d1, d0) <= TOL, maxrel(d1, d0)


@pytest.mark.parametrize("model,name", [(3, "WHITEBEAR_I"), (2, "ESFMT"), (0, "ROSENFELD")])
def test_ao_reference_force_matches_oracle_single_count(cdft, model, name):
    """ao_force=AO_FORCE_REFERENCE against the oracle's exact=False (setFundamentalMeasures_AO keeps one of
    T_jk, T_kj); the free energy is the same in both modes, the force differs where the tensor measures do."""
    rho = density(0x5EED8200 + model)
    w = oracle.fmt_weights(SHAPE, DX, HSD)
    wp = oracle.fmt_weights(SHAPE, DX, 2 * RP)
    A, B = (1.0, -0.5) if model == 2 else (1.0, 0.0)
    o = oracle.FMT(model, SHAPE, DX, HSD, weights=w, A=A, B=B)
    o.set_ao(RP, RHO_P, ao_weights=wp, exact=False)
    F0, d0 = o.eval(rho)
    s = cdft.System(SHAPE, DX, HSD, model=getattr(cdft, name), A=A, B=B, weights=w, ao_Rp=RP, ao_reservoir=RHO_P,
                    ao_weights=wp, ao_force=cdft.AO_FORCE_REFERENCE)
    F1, d1 = s.fmt_eval(rho)
    assert abs(F1 - F0) <= TOL * abs(F0), (F1, F0)
    assert maxrel(d1, d0) <= TOL, maxrel(d1, d0)
    e = cdft.System(SHAPE, DX, HSD, model=getattr(cdft, name), A=A, B=B, weights=w, ao_Rp=RP, ao_reservoir=RHO_P,
                    ao_weights=wp)
    F2, d2 = e.fmt_eval(rho)
    assert F2 == pytest.approx(F1, rel=1e-14)
    if model != 0:  # Rosenfeld has no tensor measures: both modes agree
        assert maxrel(d2, d1) > 1e3 * TOL
    else:
        assert maxrel(d2, d1) <= 1e-14


def test_full_dft_with_ao(cdft):
    kT, mu = 1.0, -0.4
    rho = density(0x5EED8100, rho0=0.4)
    z = np.arange(SHAPE[2]) * DX
    vext = np.broadcast_to(np.exp(-z / 0.3), SHAPE).copy()
    w = oracle.fmt_weights(SHAPE, DX, HSD)
    wp = oracle.fmt_weights(SHAPE, DX, 2 * RP)
    o = oracle.DFT(SHAPE, DX, HSD, model=oracle.WBI, mf=True, kT=kT, mu=mu, vext=vext)
    o.set_ao(RP, RHO_P, ao_weights=wp)
    p0, d0 = o.eval(rho)
    s = cdft.System(SHAPE, DX, HSD, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mu=mu, vext=vext, mean_field=True,
                    kT=kT, ao_Rp=RP, ao_reservoir=RHO_P, ao_weights=wp)
    p1, d1 = s.eval(rho)
    for k in range(5):
        assert abs(p1[k] - p0[k]) <= TOL * abs(p0[0])
    assert maxrel(d1, d0) <= TOL
    # batches: each member as the single system
    b = cdft.System(SHAPE, DX, HSD, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mus=[mu, mu], vexts=[vext, vext],
                    mean_field=True, kT=kT, ao_Rp=RP, ao_reservoir=RHO_P, ao_weights=wp)
    pb, db = b.eval(np.array([rho, density(5)]))
    assert np.max(np.abs(pb[0] - p1)) <= 1e-14 * abs(p1[0])
    assert maxrel(db[0], d1) <= 1e-14


def test_ao_device_weights_and_long_axis(cdft):
    """Device-generated weights at 256 along z: the force is the derivative of F (central differences)."""
    shape = (24, 16, 256)
    rho = density(0x5EED8200, shape)
    s = cdft.System(shape, 0.05, HSD, model=cdft.WHITEBEAR_I, ao_Rp=RP, ao_reservoir=RHO_P)
    F, dF = s.fmt_eval(rho)
    v = 2 * oracle.splitmix64_uniform(int(np.prod(shape)), 77).reshape(shape) - 1
    h = 1e-5
    fd = (s.fmt_eval(rho + h * v)[0] - s.fmt_eval(rho - h * v)[0]) / (2 * h)
    assert abs(fd - np.sum(dF * v)) <= 1e-7 * np.sum(np.abs(dF * v))


def spt_free_volume(rho, d, dp):
    """Free-volume fraction of a polymer of diameter dp among hard spheres of diameter d from scaled-particle
    theory (Lekkerkerker et al., Europhys. Lett. 20, 559 (1992)): (1 - eta) exp(-A g - B g^2 - C g^3),
    g = eta / (1 - eta), q = dp / d, A = 3q + 3q^2 + q^3, B = 9q^2/2 + 3q^3, C = 3q^3.  Rosenfeld's
    functional reproduces it: exp(-sum_a dPhi_a w^p_a) at uniform density."""
    eta = np.pi * rho * d ** 3 / 6
    q, g = dp / d, eta / (1 - eta)
    return (1 - eta) * np.exp(-(3 * q + 3 * q * q + q ** 3) * g - (4.5 * q * q + 3 * q ** 3) * g * g - 3 * q ** 3 * g ** 3)


def test_bulk_ao_scaled_particle_theory(cdft):
    """FMT::BulkFex / BulkMuex with the AO polymer (src/FMT.cpp:616-716) for Rosenfeld: f_AO = rho_p (1 - alpha),
    mu_AO = -rho_p d alpha / d rho (complex-step derivative of the closed form)."""
    s = cdft.System(SHAPE, DX, HSD, model=cdft.ROSENFELD, ao_Rp=RP, ao_reservoir=RHO_P)
    s0 = cdft.System(SHAPE, DX, HSD, model=cdft.ROSENFELD)
    for rho in (0.2, 0.5, 0.8):
        (f, mu), (f0, mu0) = s.bulk(rho), s0.bulk(rho)
        alpha = spt_free_volume(rho, HSD, 2 * RP)
        dalpha = spt_free_volume(rho + 1e-30j, HSD, 2 * RP).imag / 1e-30
        assert abs((f - f0) - RHO_P * (1 - alpha)) <= 1e-13 * RHO_P
        assert abs((mu - mu0) + RHO_P * dalpha) <= 1e-12 * abs(RHO_P * dalpha)


@pytest.mark.parametrize("model,name", MODELS)
def test_uniform_density_matches_bulk(cdft, model, name):
    """Uniform rho on device weights: F/V = f_ex(rho) and dF/dV = mu_ex(rho), AO terms included."""
    shape, rb = (16, 16, 16), 0.5
    s = cdft.System(shape, DX, HSD, model=getattr(cdft, name), ao_Rp=RP, ao_reservoir=RHO_P)
    F, dF = s.fmt_eval(np.full(shape, rb))
    f, mu = s.bulk(rb)
    V, dV = np.prod(shape) * DX ** 3, DX ** 3
    fex, muex = f - (rb * np.log(rb) - rb), mu - np.log(rb)
    assert abs(F / V - fex) <= 1e-11 * abs(fex), (F / V, fex)
    assert np.max(np.abs(dF / dV - muex)) <= 1e-11 * abs(muex)


def test_ao_refuses_mixed_and_hessian(cdft):
    with pytest.raises(cdft.CdftError):
        cdft.System(SHAPE, DX, HSD, precision=cdft.MIXED, ao_Rp=RP, ao_reservoir=RHO_P)
    s = cdft.System(SHAPE, DX, HSD, ao_Rp=RP, ao_reservoir=RHO_P)
    rho = density(1)
    with pytest.raises(cdft.CdftError):
        s.hessian_dot(rho, rho)
