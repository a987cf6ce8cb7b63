"""Pins of the oracle's hard-sphere mixture (orc_fmt_eval_mix, the restatement of FMT::calculateFreeEnergyAndDerivatives
over several species, src/FMT.cpp:78-111, 363-427, with the per-species collapse of include/Species.h:310-329)
against what the mathematics fixes: identical species reduce to one, a uniform mixture gives the textbook bulk
free-energy densities of the models (scaled-particle variables n0..n3) and their chemical potentials, and dF_s is
the derivative of F with respect to each species' density."""
import numpy as np
import pytest

import oracle


def bulk_phi(model, rhos, ds):
    """Bulk free-energy density of a hard-sphere mixture in the scaled-particle variables: Rosenfeld (Percus-Yevick
    compressibility), White Bear I and RSLT (Boublik-Mansoori-Carnahan-Starling-Leland), White Bear II
    (Hansen-Goos and Roth 2006).  Complex-valued so that complex-step derivatives give the chemical potentials."""
    rhos = np.asarray(rhos, dtype=complex)
    R = np.asarray(ds, dtype=float) / 2
    n0 = rhos.sum()
    n1 = (rhos * R).sum()
    n2 = (4 * np.pi * R ** 2 * rhos).sum()
    n3 = (4 * np.pi / 3 * R ** 3 * rhos).sum()
    L = np.log(1 - n3)
    if model == oracle.ROSENFELD:
        return -n0 * L + n1 * n2 / (1 - n3) + n2 ** 3 / (24 * np.pi * (1 - n3) ** 2)
    if model in (oracle.WBI, oracle.RSLT):
        return -n0 * L + n1 * n2 / (1 - n3) + n2 ** 3 * (n3 + (1 - n3) ** 2 * L) / (36 * np.pi * n3 ** 2 * (1 - n3) ** 2)
    p2 = (2 * n3 - n3 ** 2 + 2 * (1 - n3) * L) / n3
    p3 = (2 * n3 - 3 * n3 ** 2 + 2 * n3 ** 3 + 2 * (1 - n3) ** 2 * L) / n3 ** 2
    return -n0 * L + n1 * n2 * (1 + p2 / 3) / (1 - n3) + n2 ** 3 * (1 - p3 / 3) / (24 * np.pi * (1 - n3) ** 2)


def seeded(shape, mean, seed, amp):
    rng = np.random.default_rng(seed)
    return mean * (1 + amp * rng.uniform(-1, 1, shape))


@pytest.mark.parametrize("model,A,B", [(oracle.WBI, 1, -1), (oracle.ROSENFELD, 1, 0), (oracle.ESFMT, 1.3, -0.4)])
def test_identical_species_reduce_to_one(model, A, B):
    shape, dx, hsd = (12, 10, 14), 0.1, 0.6
    rho = seeded(shape, 0.4, 5, 0.1)
    w = oracle.fmt_weights(shape, dx, hsd)
    F1, dF1 = oracle.fmt_eval(model, rho, dx, hsd, weights=w, A=A, B=B)
    Fm, dFm = oracle.fmt_eval_mix(model, np.stack([0.25 * rho, 0.75 * rho]), dx, [hsd, hsd],
                                  weights=np.stack([w, w]), A=A, B=B)
    assert abs(Fm - F1) <= 1e-14 * abs(F1)
    for s in range(2):
        assert np.abs(dFm[s] - dF1).max() <= 1e-13 * np.abs(dF1).max()


@pytest.mark.parametrize("model", [oracle.ROSENFELD, oracle.RSLT, oracle.WBI, oracle.WBII])
def test_uniform_mixture_is_the_bulk_mixture(model):
    shape, dx = (20, 20, 20), 0.1
    ds = [1.0, 0.6, 0.8]
    rhos = [0.3, 0.5, 0.2]
    w = np.stack([oracle.fmt_weights(shape, dx, d, mode=oracle.WEIGHTS_ACCURATE) for d in ds])
    rho = np.stack([np.full(shape, r) for r in rhos])
    F, dF = oracle.fmt_eval_mix(model, rho, dx, ds, weights=w)
    V = np.prod(shape) * dx ** 3
    f = bulk_phi(model, rhos, ds).real
    assert abs(F / V - f) <= 1e-12 * abs(f)
    h = 1e-30
    for s in range(3):
        r = np.array(rhos, dtype=complex)
        r[s] += 1j * h
        mu = bulk_phi(model, r, ds).imag / h
        assert np.abs(dF[s] / dx ** 3 - mu).max() <= 1e-12 * abs(mu)


def test_a_second_diameter_changes_the_answer():
    """A mixture of diameters 1.0 and 0.6 is not the one-diameter system: the bulk check above would also pass for a
    restatement that used one diameter for all species only if the formula did; this one pins that it does not."""
    shape, dx = (20, 20, 20), 0.1
    w1, w2 = (oracle.fmt_weights(shape, dx, d, mode=oracle.WEIGHTS_ACCURATE) for d in (1.0, 0.6))
    rho = np.stack([np.full(shape, 0.3), np.full(shape, 0.5)])
    F, _ = oracle.fmt_eval_mix(oracle.WBI, rho, dx, [1.0, 0.6], weights=np.stack([w1, w2]))
    V = np.prod(shape) * dx ** 3
    assert abs(F / V - bulk_phi(oracle.WBI, [0.3, 0.5], [1.0, 0.6]).real) <= 1e-12 * abs(F / V)
    assert abs(F / V - bulk_phi(oracle.WBI, [0.3, 0.5], [1.0, 1.0]).real) > 1e-2 * abs(F / V)


@pytest.mark.parametrize("model,A,B", [(oracle.WBI, 1, -1), (oracle.ESFMT, 1.3, -0.4), (oracle.RSLT, 1, 0)])
def test_mixture_central_differences(model, A, B):
    shape, dx = (10, 8, 12), 0.1
    ds = [0.6, 0.45]
    w = np.stack([oracle.fmt_weights(shape, dx, d, mode=oracle.WEIGHTS_ACCURATE) for d in ds])
    rho = np.stack([seeded(shape, 0.35, 11, 0.2), seeded(shape, 0.5, 12, 0.2)])
    _, dF = oracle.fmt_eval_mix(model, rho, dx, ds, weights=w, A=A, B=B)
    rng = np.random.default_rng(3)
    h = 1e-5
    for s in range(2):
        for _ in range(3):
            i = tuple(int(rng.integers(0, n)) for n in shape)
            rp, rm = rho.copy(), rho.copy()
            rp[(s,) + i] += h
            rm[(s,) + i] -= h
            Fp, _ = oracle.fmt_eval_mix(model, rp, dx, ds, weights=w, A=A, B=B)
            Fm, _ = oracle.fmt_eval_mix(model, rm, dx, ds, weights=w, A=A, B=B)
            fd = (Fp - Fm) / (2 * h)
            assert abs(fd - dF[(s,) + i]) <= 1e-6 * abs(dF[(s,) + i]) + 1e-12


def bulk_mu_hs(model, rhos, ds):
    """Excess chemical potentials d(phi)/d(rho_s) of the bulk mixture, by complex steps."""
    h = 1e-30
    out = []
    for s in range(len(rhos)):
        r = np.array(rhos, dtype=complex)
        r[s] += 1j * h
        out.append(bulk_phi(model, r, ds).imag / h)
    return np.array(out)


def test_mixture_dft_adds_each_species_ideal_gas_field_and_mu():
    """MixDFT (the mixture's DFT::calculateFreeEnergyAndDerivatives, src/DFT.cpp:360-425): uniform densities give
    F / V = phi_hs + sum_s rho_s (ln rho_s - 1) + sum_s rho_s (V_s - mu_s) and dF_s / dV = mu_hs_s + ln rho_s
    + V_s - mu_s; a field on one species and a mu on the other show which species each term belongs to."""
    shape, dx, ds = (16, 16, 16), 0.1, [1.0, 0.6]
    rhos = [0.3, 0.5]
    mu = [0.7, -1.2]
    V1 = 0.35
    w = np.stack([oracle.fmt_weights(shape, dx, d, mode=oracle.WEIGHTS_ACCURATE) for d in ds])
    m = oracle.MixDFT(shape, dx, ds, model=oracle.WBI, weights=w, mu=mu, vext=[None, np.full(shape, V1)])
    parts, dF = m.eval(np.stack([np.full(shape, r) for r in rhos]))
    V = np.prod(shape) * dx ** 3
    ideal = sum(r * (np.log(r) - 1) for r in rhos)
    ext = rhos[1] * V1 - rhos[0] * mu[0] - rhos[1] * mu[1]
    phi = bulk_phi(oracle.WBI, rhos, ds).real
    assert abs(parts[1] / V - ideal) <= 1e-12 * abs(ideal)
    assert abs(parts[2] / V - phi) <= 1e-12 * abs(phi)
    assert abs(parts[4] / V - ext) <= 1e-12 * abs(ext)
    assert abs(parts[0] / V - (ideal + phi + ext)) <= 1e-12 * abs(ideal + phi + ext)
    mhs = bulk_mu_hs(oracle.WBI, rhos, ds)
    want = [mhs[0] + np.log(rhos[0]) - mu[0], mhs[1] + np.log(rhos[1]) + V1 - mu[1]]
    for s in range(2):
        assert np.abs(dF[s] / dx ** 3 - want[s]).max() <= 1e-11 * max(1.0, abs(want[s]))


def test_mixture_fire_finds_the_bulk_coexisting_densities():
    """FIRE2 over two species (src/Minimizer.cpp:160-327, one state over both species' aliases) from uniform
    densities away from equilibrium: with mu_s = ln rho*_s + mu_hs_s(rho*) (the bulk closed form) the minimum of
    the grand potential is the uniform mixture rho*, which the solver must reach."""
    shape, dx, ds = (8, 8, 8), 0.1, [1.0, 0.6]
    target = [0.25, 0.4]
    mhs = bulk_mu_hs(oracle.WBI, target, ds)
    mu = [np.log(target[s]) + mhs[s] for s in range(2)]
    w = np.stack([oracle.fmt_weights(shape, dx, d, mode=oracle.WEIGHTS_ACCURATE) for d in ds])
    m = oracle.MixDFT(shape, dx, ds, model=oracle.WBI, weights=w, mu=mu)
    f = oracle.Fire(m, np.stack([np.full(shape, 0.2), np.full(shape, 0.45)]), dt=1e-3, dt_max=1e-2)
    for _ in range(3000):
        rc, info = f.step()
        assert rc != 2
        if info["fmax_raw"] < 1e-10:
            break
    assert info["fmax_raw"] < 1e-10
    rho = f.state()[0]
    for s in range(2):
        assert np.abs(rho[s] / target[s] - 1).max() < 1e-8


def test_one_species_mixture_fire_is_the_single_species_fire():
    """With one species the mixture's FIRE is fireMinimizer2 on that species' DFT, step for step."""
    shape, dx, hsd = (8, 8, 8), 0.1, 0.8
    rng = np.random.default_rng(9)
    rho0 = 0.4 * (1 + 0.1 * rng.uniform(-1, 1, shape))
    vext = 0.2 * rng.uniform(-1, 1, shape)
    o = oracle.DFT(shape, dx, hsd, mu=0.3, vext=vext)
    w = oracle.fmt_weights(shape, dx, hsd)
    m = oracle.MixDFT(shape, dx, [hsd], model=oracle.WBI, weights=w[None], mu=[0.3], vext=[vext])
    f1, f2 = oracle.Fire(o, rho0), oracle.Fire(m, rho0[None])
    for _ in range(30):
        rc1, a = f1.step()
        rc2, b = f2.step()
        assert rc1 == rc2
        assert abs(a["F"] - b["F"]) <= 1e-12 * abs(a["F"])
        assert a["dt"] == b["dt"] and a["alpha"] == b["alpha"]
    assert np.abs(f1.state()[0] - f2.state()[0][0]).max() <= 1e-12 * rho0.max()


def test_mixture_species_constraints():
    """Species::setFixedMass / setFixedBackground per species of a mixture (include/Species.h:23-35,
    src/Species.cpp:64-112): the fixed-mass species is rescaled to its mass, its force projected so that
    sum dF rho = 0 with the multiplier reported as its chemical potential; the fixed-background species has zero
    force on the boundary planes; a one-species mixture with a fixed mass is the single-species DFT with it."""
    shape, dx, ds = (8, 10, 12), 0.1, [1.0, 0.6]
    V = np.prod(shape) * dx ** 3
    rho = np.stack([seeded(shape, 0.3, 21, 0.2), seeded(shape, 0.45, 22, 0.2)])
    w = np.stack([oracle.fmt_weights(shape, dx, d) for d in ds])
    mass = 0.5 * V
    m = oracle.MixDFT(shape, dx, ds, model=oracle.WBI, weights=w, mu=[0.2, 0.7], fixed_mass=[mass, 0.0],
                      boundary=[0, 1])
    parts, dF = m.eval(rho)
    mu = m.chem_potential
    free = oracle.MixDFT(shape, dx, ds, model=oracle.WBI, weights=w, mu=[0.0, 0.7])
    r0 = rho.copy()
    r0[0] *= mass / (rho[0].sum() * dx ** 3)
    p0, d0 = free.eval(r0)  # unconstrained force at the rescaled density, mu_0 = 0
    lam = (d0[0] * r0[0]).sum() / mass
    assert abs(mu[0] - lam) <= 1e-12 * abs(lam) and mu[1] == 0.7
    assert np.abs(dF[0] - (d0[0] - lam * dx ** 3)).max() <= 1e-12 * np.abs(d0[0]).max()
    assert abs((dF[0] * r0[0]).sum()) <= 1e-10 * np.abs(d0[0] * r0[0]).sum()
    bnd = np.zeros(shape, dtype=bool)
    bnd[0, :, :] = bnd[:, 0, :] = bnd[:, :, 0] = True
    assert np.all(dF[1][bnd] == 0.0)
    assert np.abs(dF[1][~bnd] - d0[1][~bnd]).max() <= 1e-12 * np.abs(d0[1]).max()
    assert abs(parts[0] - p0[0]) <= 1e-12 * abs(p0[0])
    # one species with a fixed mass: the single-species DFT
    o = oracle.DFT(shape, dx, ds[0], mu=0.2)
    o.set_constraints(fixed_mass=mass)
    p1, d1 = o.eval(rho[0])
    m1 = oracle.MixDFT(shape, dx, ds[:1], model=oracle.WBI, weights=w[:1], mu=[0.2], fixed_mass=[mass])
    p2, d2 = m1.eval(rho[:1])
    assert abs(p2[0] - p1[0]) <= 1e-12 * abs(p1[0])
    assert np.abs(d2[0] - d1).max() <= 1e-12 * np.abs(d1).max()
    assert abs(m1.chem_potential[0] - o.chem_potential) <= 1e-12 * abs(o.chem_potential)
