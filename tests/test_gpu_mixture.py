"""GPU parity of the hard-sphere mixture (cdft_mixture_*) against the CPU oracle's mixture restatement (FMT over
several species, src/FMT.cpp:78-111, 363-427), through the C ABI.  Both sides take the same cell weights per
species, so the comparison isolates the pipeline; FP64 tolerance 1e-11 relative in F and max-norm dF_s."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-11

MODELS = [("ROSENFELD", oracle.ROSENFELD, 1.0, 0.0), ("RSLT", oracle.RSLT, 1.0, 0.0),
          ("ESFMT", oracle.ESFMT, 1.3, -0.4), ("WHITEBEAR_I", oracle.WBI, 1.0, -1.0),
          ("WHITEBEAR_II", oracle.WBII, 1.0, -1.0)]


def seeded(shape, mean, seed, amp=0.1):
    u = oracle.splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    return mean * (1.0 + amp * (2.0 * u - 1.0))


def rel(a, b):
    return abs(a - b) / abs(b)


def maxrel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("name,kind,A,B", MODELS)
def test_two_species_match_the_oracle(cdft, name, kind, A, B):
    shape, dx, ds = (24, 16, 32), 0.1, [1.0, 0.6]
    w = np.stack([oracle.fmt_weights(shape, dx, d) for d in ds])
    rho = np.stack([seeded(shape, 0.3, 0x5EED7001), seeded(shape, 0.5, 0x5EED7002)])
    F0, dF0 = oracle.fmt_eval_mix(kind, rho, dx, ds, weights=w, A=A, B=B)
    mix = cdft.Mixture(shape, dx, ds, model=getattr(cdft, name), weights=list(w), A=A, B=B)
    F1, dF1 = mix.fmt_eval(rho)
    mix.close()
    print(name, "F", rel(F1, F0), "dF", [maxrel(dF1[s], dF0[s]) for s in range(2)])
    assert rel(F1, F0) <= TOL
    for s in range(2):
        assert maxrel(dF1[s], dF0[s]) <= TOL


def test_three_species_with_device_weights(cdft):
    """Device-generated weights (double-double cell integrals) against the oracle's IEEE-quad weights."""
    shape, dx, ds = (32, 24, 48), 0.05, [1.0, 0.6, 0.8]
    w = np.stack([oracle.fmt_weights(shape, dx, d, mode=oracle.WEIGHTS_ACCURATE) for d in ds])
    rho = np.stack([seeded(shape, 0.25, 0x5EED7011), seeded(shape, 0.4, 0x5EED7012), seeded(shape, 0.2, 0x5EED7013)])
    F0, dF0 = oracle.fmt_eval_mix(oracle.WBI, rho, dx, ds, weights=w)
    mix = cdft.Mixture(shape, dx, ds, model=cdft.WHITEBEAR_I)
    F1, dF1 = mix.fmt_eval(rho)
    mix.close()
    print("3 species F", rel(F1, F0), "dF", [maxrel(dF1[s], dF0[s]) for s in range(3)])
    assert rel(F1, F0) <= TOL
    for s in range(3):
        assert maxrel(dF1[s], dF0[s]) <= TOL


def test_one_species_mixture_is_the_single_species_system(cdft):
    shape, dx, hsd = (32, 32, 32), 0.1, 1.0
    rho = seeded(shape, 0.6, 0x5EED0001, 0.05)
    sysm = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I)
    F0, dF0 = sysm.fmt_eval(rho)
    sysm.close()
    mix = cdft.Mixture(shape, dx, [hsd], model=cdft.WHITEBEAR_I)
    F1, dF1 = mix.fmt_eval(rho[None])
    mix.close()
    assert rel(F1, F0) <= 1e-13
    assert maxrel(dF1[0], dF0) <= 1e-13


def test_accumulate_and_device_pointers(cdft):
    shape, dx, ds = (16, 24, 32), 0.1, [0.8, 0.5]
    rho = np.stack([seeded(shape, 0.35, 0x5EED7021), seeded(shape, 0.45, 0x5EED7022)])
    mix = cdft.Mixture(shape, dx, ds, model=cdft.WHITEBEAR_II)
    F0, dF0 = mix.fmt_eval(rho)
    base = np.stack([np.full(shape, 1.5), np.full(shape, -0.25)])
    acc = base.copy()
    F1, out = mix.fmt_eval(rho, dF=acc, accumulate=True)
    assert out is not None and F1 == F0
    assert np.abs(acc - (base + dF0)).max() <= 1e-15 * np.abs(base + dF0).max()
    rd = [cdft.DeviceArray.from_host(rho[s]) for s in range(2)]
    fd = [cdft.DeviceArray(shape) for _ in range(2)]
    F2 = mix.fmt_eval_device([a.ptr for a in rd], [a.ptr for a in fd])
    assert F2 == F0
    for s in range(2):
        assert np.array_equal(fd[s].download(), dF0[s])
    mix.close()


def test_eta_too_large_is_reported(cdft):
    shape, dx, ds = (16, 16, 16), 0.1, [1.0, 0.6]
    mix = cdft.Mixture(shape, dx, ds, model=cdft.WHITEBEAR_I)
    rho = np.stack([np.full(shape, 1.5), np.full(shape, 2.0)])
    with pytest.raises(cdft.EtaTooLarge):
        mix.fmt_eval(rho)
    ok = np.stack([np.full(shape, 0.3), np.full(shape, 0.3)])
    F, _ = mix.fmt_eval(ok)  # the flag was cleared
    assert np.isfinite(F)
    mix.close()


def close(a, b, tol):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300)


def test_mixture_dft_eval_matches_the_oracle(cdft):
    """cdft_mixture_eval (DFT::calculateFreeEnergyAndDerivatives over several species, src/DFT.cpp:360-425): per
    species ideal gas, V_ext and -mu N beside the mixture's FMT; a field on one species only."""
    shape, dx, ds = (16, 24, 32), 0.1, [1.0, 0.6]
    w = np.stack([oracle.fmt_weights(shape, dx, d) for d in ds])
    rho = np.stack([seeded(shape, 0.3, 0x5EED7031), seeded(shape, 0.45, 0x5EED7032)])
    v1 = 0.3 * (2 * oracle.splitmix64_uniform(int(np.prod(shape)), 0x5EED7033).reshape(shape) - 1)
    mu = [0.4, -0.9]
    o = oracle.MixDFT(shape, dx, ds, model=oracle.WBI, weights=w, mu=mu, vext=[None, v1])
    p0, dF0 = o.eval(rho)
    mix = cdft.Mixture(shape, dx, ds, model=cdft.WHITEBEAR_I, weights=list(w), mu=mu, vext=[None, v1])
    p1, dF1 = mix.eval(rho)
    print("mixture DFT parts", p0, p1)
    for k in (0, 1, 2, 4):
        assert close(p1[k], p0[k], TOL), (k, p1[k], p0[k])
    for s in range(2):
        assert maxrel(dF1[s], dF0[s]) <= TOL
    # only_fex: the ideal-gas terms leave, field and mu stay
    p2, dF2 = mix.eval(rho, only_fex=True)
    oi = oracle.MixDFT(shape, dx, ds, model=oracle.WBI, weights=w, ideal=False, mu=mu, vext=[None, v1])
    p3, dF3 = oi.eval(rho)
    assert p2[1] == 0.0 and close(p2[0], p3[0], TOL)
    for s in range(2):
        assert maxrel(dF2[s], dF3[s]) <= TOL
    mix.close()


def test_mixture_fire_replays_the_oracle(cdft):
    """cdft_mixture_fire_create: one FIRE2 state over both species (src/Minimizer.cpp:160-327), step for step
    against the oracle's FIRE over the mixture."""
    shape, dx, ds = (16, 16, 16), 0.1, [1.0, 0.6]
    w = np.stack([oracle.fmt_weights(shape, dx, d) for d in ds])
    rho0 = np.stack([seeded(shape, 0.3, 0x5EED7041), seeded(shape, 0.4, 0x5EED7042)])
    v0 = 0.2 * (2 * oracle.splitmix64_uniform(int(np.prod(shape)), 0x5EED7043).reshape(shape) - 1)
    mu = [0.5, 0.2]
    o = oracle.MixDFT(shape, dx, ds, model=oracle.WBI, weights=w, mu=mu, vext=[v0, None])
    mix = cdft.Mixture(shape, dx, ds, model=cdft.WHITEBEAR_I, weights=list(w), mu=mu, vext=[v0, None])
    fo = oracle.Fire(o, rho0)
    fg = cdft.Fire(mix, rho0)
    F0 = o.eval(rho0)[0][0]
    for i in range(40):
        rc0, a = fo.step()
        rc1, b = fg.step()
        assert rc0 == rc1, (i, rc0, rc1)
        assert close(a["F"], b["F"], 1e-10), (i, a["F"], b["F"])
        assert close(a["dt"], b["dt"], 1e-12) and close(a["alpha"], b["alpha"], 1e-12)
        assert close(a["fmax"], b["fmax_alias"], 1e-9) and close(a["fmax_raw"], b["fmax_raw"], 1e-9)
        # fireMinimizer2::get_convergence_monitor = |vv| / V, the lattice volume (include/Minimizer.h:105)
        assert close(b["monitor"], abs(a["vv"]) / (np.prod(shape) * dx ** 3), 1e-9)
    r0 = fo.state()[0]
    r1 = fg.state()[0]
    assert r1.shape == (2,) + shape
    assert np.abs(r0 - r1).max() <= 1e-10 * np.abs(r0).max()
    assert b["F"] < F0
    fg.close()
    mix.close()


def test_mixture_fire_reaches_the_bulk_mixture(cdft):
    """With mu_s = ln rho*_s + mu_hs_s(rho*) from the bulk White Bear closed form, FIRE over the mixture from other
    uniform densities ends at rho*."""
    from test_oracle_mixture import bulk_mu_hs
    shape, dx, ds = (16, 16, 16), 0.1, [1.0, 0.6]
    target = [0.25, 0.4]
    mhs = bulk_mu_hs(oracle.WBI, target, ds)
    mu = [np.log(target[s]) + mhs[s] for s in range(2)]
    mix = cdft.Mixture(shape, dx, ds, model=cdft.WHITEBEAR_I, mu=mu)
    f = cdft.Fire(mix, np.stack([np.full(shape, 0.2), np.full(shape, 0.45)]))
    rc, rep = f.run(5000, 1e-10)
    assert rc != 2 and rep["fmax_raw"] <= 1e-10
    rho = f.state()[0]
    for s in range(2):
        assert np.abs(rho[s] / target[s] - 1).max() < 1e-7
    f.close()
    mix.close()


def test_mixture_species_constraints_match_the_oracle(cdft):
    """Fixed mass on one species and a fixed background on the other (Species::setFixedMass /
    setFixedBackground per species of a mixture): evaluation, multiplier, and a FIRE replay with the constraints."""
    shape, dx, ds = (16, 16, 24), 0.1, [1.0, 0.6]
    V = np.prod(shape) * dx ** 3
    w = np.stack([oracle.fmt_weights(shape, dx, d) for d in ds])
    rho = np.stack([seeded(shape, 0.3, 0x5EED7051), seeded(shape, 0.4, 0x5EED7052)])
    kw = dict(mu=[0.3, 0.5], fixed_mass=[0.0, 0.42 * V], boundary=[cdft.FIXED_BACKGROUND, 0])
    o = oracle.MixDFT(shape, dx, ds, model=oracle.WBI, weights=w, **kw)
    mix = cdft.Mixture(shape, dx, ds, model=cdft.WHITEBEAR_I, weights=list(w), **kw)
    p0, dF0 = o.eval(rho)
    p1, dF1 = mix.eval(rho)
    for k in (0, 1, 2, 4):
        assert close(p1[k], p0[k], TOL), (k, p1[k], p0[k])
    for s in range(2):
        assert maxrel(dF1[s], dF0[s]) <= TOL
    mu0, mu1 = o.chem_potential, mix.chem_potential()
    assert mu1[0] == 0.3 and close(mu1[1], mu0[1], TOL)
    fo = oracle.Fire(o, rho)
    fg = cdft.Fire(mix, rho)
    for i in range(20):
        rc0, a = fo.step()
        rc1, b = fg.step()
        assert rc0 == rc1, (i, rc0, rc1)
        assert close(a["F"], b["F"], 1e-10), (i, a["F"], b["F"])
        assert close(a["dt"], b["dt"], 1e-12)
    r0, r1 = fo.state()[0], fg.state()[0]
    assert np.abs(r0 - r1).max() <= 1e-10 * np.abs(r0).max()
    assert abs(r1[1].sum() * dx ** 3 / (0.42 * V) - 1) <= 1e-12
    assert close(mix.chem_potential()[1], o.chem_potential[1], 1e-9)
    fg.close()
    mix.close()


def test_mixture_x_slab_section_matches(cdft, monkeypatch):
    """The measures' y / z / y section run x-slab by x-slab (a 1 MB cache budget forces 7 slabs at this size) gives
    the forces of the whole-array section bit for bit, F to rounding, and the oracle's mixture to 1e-11."""
    shape, dx, ds = (32, 24, 48), 0.05, [1.0, 0.6, 0.8]
    rho = np.stack([seeded(shape, 0.25, 0x5EED7061), seeded(shape, 0.4, 0x5EED7062), seeded(shape, 0.2, 0x5EED7063)])
    whole = cdft.Mixture(shape, dx, ds, model=cdft.WHITEBEAR_II)
    F0, dF0 = whole.fmt_eval(rho)
    whole.close()
    monkeypatch.setenv("CDFT_SLAB_MB", "1")
    slabs = cdft.Mixture(shape, dx, ds, model=cdft.WHITEBEAR_II)
    F1, dF1 = slabs.fmt_eval(rho)
    slabs.close()
    assert np.array_equal(dF1, dF0)
    assert abs(F1 - F0) <= 1e-14 * abs(F0)
    w = np.stack([oracle.fmt_weights(shape, dx, d, mode=oracle.WEIGHTS_ACCURATE) for d in ds])
    F2, dF2 = oracle.fmt_eval_mix(oracle.WBII, rho, dx, ds, weights=w)
    assert rel(F1, F2) <= TOL
    for s in range(3):
        assert maxrel(dF1[s], dF2[s]) <= TOL


def test_mixture_with_and_without_the_trace_identity(cdft):
    """Weights that satisfy w_Txx + w_Tyy + w_Tzz = w_s (device-generated) for one species and weights that do
    not (the reference generator's, given by the caller) for the other: the mixture carries the T_zz measure and
    builds the first species' T_zz weight from its other weights; 1e-11 against the oracle."""
    shape, dx, ds = (24, 16, 32), 0.1, [1.0, 0.6]
    wf = oracle.fmt_weights(shape, dx, ds[1])  # the reference generator's rounding
    wa = oracle.fmt_weights(shape, dx, ds[0], mode=oracle.WEIGHTS_ACCURATE)
    rho = np.stack([seeded(shape, 0.3, 0x5EED7071), seeded(shape, 0.45, 0x5EED7072)])
    F0, dF0 = oracle.fmt_eval_mix(oracle.ESFMT, rho, dx, ds, weights=np.stack([wa, wf]), A=1.3, B=-0.4)
    mix = cdft.Mixture(shape, dx, ds, model=cdft.ESFMT, weights=[None, wf], A=1.3, B=-0.4)
    F1, dF1 = mix.fmt_eval(rho)
    mix.close()
    assert rel(F1, F0) <= TOL
    for s in range(2):
        assert maxrel(dF1[s], dF0[s]) <= TOL
