"""The reference-named C++ API (include/classicaldft/classicaldft.h) end to end on the GPU.

Runs classicaldft_amd/dropin_demo (built by build()) and checks its printed
numbers against the oracle: config-1 F_hs with device weights vs the oracle
on accurate weights (1e-11), bulk mu, 20 FIRE2 steps and 3 DDFT steps of a
16^3 ideal + White Bear I + LJ/QF system (1e-10), and the LJ hard-sphere
diameter against an independent scipy quadrature.
"""
import os
import subprocess

import numpy as np
import pytest
from scipy import integrate

import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def seeded(shape, rho_b, seed, amp):
    u = oracle.splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    return rho_b * (1 + amp * (2 * u - 1))


@pytest.fixture(scope="module")
def demo(cdft, tmp_path_factory):
    exe = os.path.join(ROOT, "classicaldft_amd", "dropin_demo")
    wd = tmp_path_factory.mktemp("demo")  # fireMinimizer2 writes snapshot_s0.dat to the working directory
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120, cwd=wd).stdout
    vals = {"_dir": wd}
    for line in out.splitlines():
        k, *v = line.split()
        vals[k] = [float(x) for x in v]
    return vals


def rel(a, b):
    return abs(a - b) / abs(b)


def test_config1_through_cpp_api(demo):
    shape, dx, hsd = (32, 32, 32), 0.1, 1.0
    rho = seeded(shape, 0.6, 0x5EED0001, 0.05)
    w = oracle.fmt_weights(shape, dx, hsd, mode=oracle.WEIGHTS_ACCURATE)
    F, dF = oracle.fmt_eval(oracle.WBI, rho, dx, hsd, weights=w)
    assert rel(demo["F_hs"][0], F) <= 1e-11
    assert abs(demo["dF_sum"][0] - dF.sum()) <= 1e-11 * np.abs(dF).sum()


def test_bulk_mu_through_cpp_api(demo):
    o = oracle.DFT((16, 16, 16), 0.1, 1.0, mf=True, kT=1.0)
    _, muex = oracle.bulk(oracle.WBI, 0.5, 1.0)
    assert rel(demo["mu"][0], np.log(0.5) + muex + o.a_vdw * 0.5) <= 1e-12


def test_fire_and_ddft_through_cpp_api(demo):
    shape, dx = (16, 16, 16), 0.1
    o = oracle.DFT(shape, dx, 1.0, mf=True, kT=1.0, mu=0.5, accurate_weights=True)
    f = oracle.Fire(o, seeded(shape, 0.5, 11, 0.1))
    for _ in range(20):
        rc, info = f.step()
    assert rel(demo["fire"][0], info["F"]) <= 1e-10
    assert rel(demo["fire"][1], info["fmax_raw"]) <= 1e-8
    o3 = oracle.DFT(shape, dx, 1.0, mf=True, kT=1.0, accurate_weights=True)
    rho3 = seeded(shape, 0.5, 21, 0.1)
    d = oracle.DDFT(o3, rho3)
    for _ in range(3):
        r = d.step()
    assert rel(demo["ddft"][0], r["F"]) <= 1e-10
    assert rel(demo["ddft"][1], rho3.sum() * dx ** 3) <= 1e-12


def test_lj_hard_sphere_diameter(demo):
    kT, rc = 0.8, 2.5
    rmin = 2 ** (1 / 6)
    shift = 4 * (rc ** -12 - rc ** -6)

    def V(r):
        return 4 * (r ** -12 - r ** -6) - shift

    vmin = V(rmin)
    val, _ = integrate.quad(lambda r: 1 - np.exp(-(V(r) - vmin) / kT) if r > 0 else 1.0, 0, rmin, epsabs=1e-14,
                            epsrel=1e-13, limit=200)
    assert abs(demo["hsd_lj_kT0.8"][0] - val) < 1e-11


def test_hessian_through_cpp_api(demo):
    shape, dx = (16, 16, 16), 0.1
    rho = seeded(shape, 0.5, 41, 0.1)
    v = 2 * oracle.splitmix64_uniform(int(np.prod(shape)), 43).reshape(shape) - 1
    o = oracle.DFT(shape, dx, 1.0, mf=True, kT=1.0, accurate_weights=True)
    Hv = o.hess(rho, v)
    assert abs(demo["hess"][0] - Hv.sum()) <= 1e-10 * np.abs(Hv).sum()
    assert rel(demo["hess"][1], np.sum(v * Hv)) <= 1e-10
    of = oracle.DFT(shape, dx, 1.0, ideal=False, accurate_weights=True)
    Hf = of.hess(rho, v)
    assert abs(demo["hess_fmt"][0] - Hf.sum()) <= 1e-10 * np.abs(Hf).sum()
    assert rel(demo["hess_fmt"][1], np.sum(v * Hf)) <= 1e-10


def test_fire_snapshot_file(demo):
    """fireMinimizer2 wrote snapshot_s0.dat (src/Minimizer.cpp:259-268) in the reference's density format."""
    from classicaldft_amd import formats
    rho = formats.read_density(os.path.join(demo["_dir"], "snapshot_s0.dat"), (16, 16, 16))
    assert np.all(np.isfinite(rho)) and np.all(rho > 0)


def test_other_potentials_through_cpp_api(demo):
    """tWF / WHDF potentials, the BH split and the Gauss_F / LE kernel rules through the reference-named classes."""
    w, a = oracle.mf_kernel2((16, 16, 16), 0.1, kind=oracle.POT_TWF, rcut=2.0, alpha=50.0, bh=True,
                             scheme=oracle.MF_GAUSS_F, ngauss=3, kT=1.2)
    _, muex = oracle.bulk(oracle.WBI, 0.5, 1.0)
    assert rel(demo["mu_twf_bh_gauss_f"][0], np.log(0.5) + muex + a * 0.5) <= 1e-12
    _, a_le = oracle.mf_kernel2((16, 16, 16), 0.1, kind=oracle.POT_WHDF, rcut=2.0, scheme=oracle.MF_LE, kT=0.9)
    assert rel(demo["avdw_whdf_le"][0], a_le) <= 1e-12


def test_ao_species_through_cpp_api(demo, cdft):
    """FMT_AO_Species (include/Species.h:380-461) with White Bear I through the reference-named classes."""
    shape, dx = (16, 16, 16), 0.1
    rho = seeded(shape, 0.45, 51, 0.1)
    o = oracle.FMT(oracle.WBI, shape, dx, 1.0, weights=oracle.fmt_weights(shape, dx, 1.0, mode=oracle.WEIGHTS_ACCURATE))
    o.set_ao(0.3, 0.4, ao_weights=oracle.fmt_weights(shape, dx, 0.6, mode=oracle.WEIGHTS_ACCURATE))
    F, dF = o.eval(rho)
    assert rel(demo["ao"][0], F) <= 1e-11
    assert abs(demo["ao"][1] - dF.sum()) <= 1e-11 * np.abs(dF).sum()
    _, mu = cdft.System(shape, dx, 1.0, ao_Rp=0.3, ao_reservoir=0.4).bulk(0.5)
    assert rel(demo["ao"][2], mu - np.log(0.5)) <= 1e-13


def test_species_constraints_through_cpp_api(demo):
    """Species::setFixedMass (evaluation and fireMinimizer2) and setFixedBackground with onlyFex through the
    reference-named classes (src/Species.cpp:64-112, src/DFT.cpp:360-425)."""
    shape, dx = (16, 16, 16), 0.1
    rho = seeded(shape, 0.5, 61, 0.1)
    M = 0.95 * rho.sum() * dx ** 3
    o = oracle.DFT(shape, dx, 1.0, mf=True, kT=1.0, accurate_weights=True)
    o.set_constraints(fixed_mass=M)
    p, dF = o.eval(rho)
    F, mu, dsum = demo["fixed_mass_eval"]
    assert rel(F, p[0]) <= 1e-11 and rel(mu, o.chem_potential) <= 1e-11
    assert abs(dsum - dF.sum()) <= 1e-11 * np.abs(dF).sum()
    f = oracle.Fire(o, o.density)
    for _ in range(10):
        _, info = f.step()
    F, mu, N = demo["fixed_mass_fire"]
    assert rel(F, info["F"]) <= 1e-10
    assert rel(mu, o.chem_potential) <= 1e-9 and rel(N, M) <= 1e-12
    rho8 = seeded(shape, 0.5, 71, 0.1)
    o8 = oracle.DFT(shape, dx, 1.0, mf=True, kT=1.0, mu=0.3, accurate_weights=True)
    o8.set_constraints(fixed_background=True)
    p8, d8 = o8.eval(rho8)
    F8, dsum8, F8x = demo["fixed_bg"]
    assert rel(F8, p8[0]) <= 1e-11 and abs(dsum8 - d8.sum()) <= 1e-11 * np.abs(d8).sum()
    ox = oracle.DFT(shape, dx, 1.0, mf=True, kT=1.0, mu=0.3, ideal=False, accurate_weights=True)
    assert rel(F8x, ox.eval(rho8)[0][0]) <= 1e-11


def test_fire_fixed_direction_through_cpp_api(demo):
    """Minimizer::set_fixed_direction (src/Minimizer.cpp:103-110) through fireMinimizer2."""
    shape, dx = (16, 16, 16), 0.1
    rho = seeded(shape, 0.5, 81, 0.1)
    e = 2 * oracle.splitmix64_uniform(int(np.prod(shape)), 83).reshape(shape) - 1
    e /= np.sqrt(np.sum(e * e))
    o = oracle.DFT(shape, dx, 1.0, mu=0.4, accurate_weights=True)
    f = oracle.Fire(o, rho)
    f.set_fixed_direction(e)
    f.reset(rho)
    for _ in range(5):
        _, info = f.step()
    assert rel(demo["fire_dir"][0], info["F"]) <= 1e-10


def test_open_ddft_through_cpp_api(demo):
    """DDFT_IF_Open (mduran/src/DDFT_IF_Open.cpp:25-463) through the reference-named class."""
    shape, dx = (16, 16, 16), 0.1
    rho = seeded(shape, 0.5, 91, 0.1)
    o = oracle.DFT(shape, dx, 1.0, mf=True, kT=1.0, accurate_weights=True)
    d = oracle.OpenDDFT(o, rho, 0.5)
    d.step()
    info = d.step()
    r = d.density()
    F, r0, r8 = demo["open_ddft"]
    assert rel(F, info["F"]) <= 1e-10 and r0 == 0.5 and rel(r8, r[8, 8, 8]) <= 1e-10


def test_points_interaction_through_cpp_api(demo):
    """Interaction(s1, s2, v, kT, pointsFile) (src/Interaction.cpp:566-860): the kernel generated on the GPU, the
    reference-format cache written and read back by a second object, and a DFT with it (cache precision: the
    reference's 6 significant digits)."""
    from classicaldft_amd import formats
    shape, dx = (16, 16, 16), 0.1
    pts = formats.read_points_file(os.path.join(demo["_dir"], "sphere_points.dat"))
    w = oracle.mf_kernel_points(shape, dx, pts)
    a1, a2, cached, F = demo["points"]
    assert rel(a1, w.sum()) <= 1e-12 and cached == 1 and rel(a2, a1) <= 1e-5
    ident = formats.potential_identifier(0, 1.0, 1.0, 2.5)
    path = os.path.join(demo["_dir"], formats.weight_cache_name(7, 7, shape, ident))
    wc = formats.read_weight_cache(path, shape, dx, ident, "sphere_points.dat")
    assert wc is not None and np.max(np.abs(wc - w)) <= 5e-6 * np.abs(w).max()
    rho = seeded(shape, 0.5, 101, 0.1)
    o = oracle.DFT(shape, dx, 1.0, accurate_weights=True)
    o.set_mean_field(wc, wc.sum())
    assert rel(F, o.eval(rho)[0][0]) <= 1e-11


def test_scaled_interaction_through_cpp_api(demo):
    """Interaction_Base::scale_interaction (src/Interaction.cpp:54-59), getW and the bulk mu that follows."""
    w, a = oracle.mf_kernel((16, 16, 16), 0.1, kT=1.0)
    a_s, w0, mu = demo["scaled_mf"]
    assert rel(a_s, 0.5 * a) <= 1e-12 and rel(w0, 0.5 * w[0, 0, 0]) <= 1e-13
    _, muex = oracle.bulk(oracle.WBI, 0.5, 1.0)
    assert rel(mu, np.log(0.5) + muex + 0.5 * a * 0.5) <= 1e-12


def test_log_det_through_cpp_api(demo):
    """Log_Det(dft, 20, 0.05, 5e-4).calculate_log_det(7, 2, false, variance) (src/Log_Det.cpp:96-166)."""
    shape, dx = (16, 16, 16), 0.1
    rho = seeded(shape, 0.5, 41, 0.1)
    o = oracle.DFT(shape, dx, 1.0, accurate_weights=True)
    ld, var = o.log_det(rho, 20, 0.05, 5e-4, 7, 2)
    assert rel(demo["log_det"][0], ld) <= 1e-10
    q = ld + rho.size * np.log(1.0 / (0.05 + 5e-4))  # the mean of v.u over the probes
    assert abs(demo["log_det"][1] ** 2 - var ** 2) <= 1e-10 * q * q
    ld2, var2 = o.log_det(rho, 12, 40.0, 5e-3, 7, 2, has_zero=True, dynamic=True)
    assert rel(demo["log_det_ddft"][0], ld2) <= 1e-10
    v = 2 * oracle.splitmix64_uniform(int(np.prod(shape)), 43).reshape(shape) - 1
    gv = o.dynamic_dot(rho, v)
    assert abs(demo["ddft_dot"][0]) <= 1e-10 * np.abs(gv).sum()  # g conserves mass: the entries sum to zero
    assert rel(demo["ddft_dot"][1], np.sum(v * gv)) <= 1e-10


def test_eigenvalues_through_cpp_api(demo):
    """Eigenvalues::calculate_eigenvector (src/Eigenvalues.cpp:72-99): the smallest and, with change_sign, the largest
    eigenvalue of the 8^3 Hessian against numpy's eigvalsh of the dense matrix built from oracle products."""
    shape, dx = (8, 8, 8), 0.2
    rho = seeded(shape, 0.4, 47, 0.1)
    o = oracle.DFT(shape, dx, 1.0, accurate_weights=True)
    N = rho.size
    H = np.stack([o.hess(rho, np.eye(N)[j].reshape(shape)).ravel() for j in range(N)], axis=1)
    ev = np.linalg.eigvalsh(0.5 * (H + H.T))
    lo, hi = demo["eigen"][:2]
    assert rel(lo, ev[0]) <= 1e-6 and rel(hi, ev[-1]) <= 1e-6


def test_coexistence_through_cpp_api(demo, cdft):
    """DFT::findSpinodal / findCoex / find_density_from_mu / getCriticalPoint (src/DFT_Coex.cpp:39-196) through the
    C++ API, checked by what defines them: the spinodal densities are stationary points of the pressure, the
    coexisting densities have equal mu and equal pressure, mu(x) = mu(0.3) gives 0.3, and (xc, Tc) is the fixed point
    of the critical-point estimate."""
    xs1, xs2, x1, x2, xmu, xc, Tc = demo["coex"]
    s = cdft.System((16, 16, 16), 0.1, 1.0, model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=0.8)

    def mu(x):
        return s.bulk(x)[1]

    def p(x):
        f, m = s.bulk(x)
        return x * m - f

    assert 0 < x1 < xs1 < xs2 < x2 < 1.2
    assert abs(mu(x1) - mu(x2)) <= 1e-7 * abs(mu(x1))
    # findCoex stops on |y2 - y1| <= tol and takes the liquid density of the last bisection step: with tol = 1e-10
    # the pressures agree to about 1e-7 relative
    assert abs(p(x1) - p(x2)) <= 1e-6 * abs(p(x1))
    h = 1e-6
    slope = abs(p(x1 + h) - p(x1 - h)) / (2 * h)
    for xs in (xs1, xs2):
        assert abs(p(xs + h) - p(xs - h)) / (2 * h) <= 1e-4 * slope
    assert abs(xmu - 0.3) <= 1e-9
    d = cdft.lj_hsd(Tc)
    # the loop ends when Tc moves by at most 1e-8 relative, xc holds the diameter of the step before
    assert abs(xc - 0.13044 * (6 / np.pi) / d ** 3) <= 1e-7 * xc
    assert abs(Tc + 0.090082 * 2 * cdft.potential_vdw(Tc) / d ** 3 * Tc) <= 1e-6 * Tc


def test_mixture_through_cpp_api(demo):
    """FMT::calculateFreeEnergyAndDerivatives with two FMT_Species (diameters 1 and 0.6) runs the device mixture;
    F and each species' force sum against the oracle's mixture on accurate weights."""
    shape, dx, ds = (16, 16, 24), 0.1, [1.0, 0.6]
    rho = np.stack([seeded(shape, 0.3, 61, 0.1), seeded(shape, 0.5, 62, 0.1)])
    w = np.stack([oracle.fmt_weights(shape, dx, d, mode=oracle.WEIGHTS_ACCURATE) for d in ds])
    F, dF = oracle.fmt_eval_mix(oracle.WBI, rho, dx, ds, weights=w)
    got = demo["mix"]
    assert rel(got[0], F) <= 1e-11
    for s in range(2):
        assert abs(got[1 + s] - dF[s].sum()) <= 1e-11 * np.abs(dF[s]).sum()


def test_mixture_dft_and_fire_through_cpp_api(demo):
    """DFT with two FMT_Species (DFT::calculateFreeEnergyAndDerivatives over several species, src/DFT.cpp:360-425)
    and fireMinimizer2 over both (src/Minimizer.cpp:160-327) against the oracle's mixture DFT and FIRE."""
    shape, dx, ds = (16, 16, 24), 0.1, [1.0, 0.6]
    rho = np.stack([seeded(shape, 0.3, 63, 0.1), seeded(shape, 0.4, 64, 0.1)])
    field = 0.2 * (2 * oracle.splitmix64_uniform(int(np.prod(shape)), 65).reshape(shape) - 1)
    w = np.stack([oracle.fmt_weights(shape, dx, d, mode=oracle.WEIGHTS_ACCURATE) for d in ds])
    o = oracle.MixDFT(shape, dx, ds, model=oracle.WBI, weights=w, mu=[0.5, 0.2], vext=[field, None])
    parts, dF = o.eval(rho)
    got = demo["mix_dft"]
    assert rel(got[0], parts[0]) <= 1e-11 and rel(got[1], parts[1]) <= 1e-12
    for s in range(2):
        assert abs(got[2 + s] - dF[s].sum()) <= 1e-11 * np.abs(dF[s]).sum()
    f = oracle.Fire(o, rho)
    for _ in range(20):
        rc, info = f.step()
    got = demo["mix_fire"]
    assert rel(got[0], info["F"]) <= 1e-10
    assert rel(got[1], info["fmax_raw"]) <= 1e-8
    r = f.state()[0]
    for s in range(2):
        assert rel(got[2 + s], r[s].sum() * dx ** 3) <= 1e-11


def test_mixture_fixed_mass_through_cpp_api(demo):
    """Species::setFixedMass on one species of a two-species DFT: F, the multiplier recorded as that species'
    chemical potential and its mass, after one evaluation and after 10 fireMinimizer2 steps, against the oracle."""
    shape, dx, ds = (16, 16, 24), 0.1, [1.0, 0.6]
    V = np.prod(shape) * dx ** 3
    rho = np.stack([seeded(shape, 0.3, 66, 0.1), seeded(shape, 0.4, 67, 0.1)])
    w = np.stack([oracle.fmt_weights(shape, dx, d, mode=oracle.WEIGHTS_ACCURATE) for d in ds])
    o = oracle.MixDFT(shape, dx, ds, model=oracle.WBI, weights=w, mu=[0.3, 0.5], fixed_mass=[0.0, 0.42 * V])
    parts, _ = o.eval(rho)
    got = demo["mix_mass"]
    assert rel(got[0], parts[0]) <= 1e-11
    assert rel(got[1], o.chem_potential[1]) <= 1e-10
    assert rel(got[2], 0.42 * V) <= 1e-12
    f = oracle.Fire(o, rho)
    for _ in range(10):
        rc, info = f.step()
    got = demo["mix_mass_fire"]
    assert rel(got[0], info["F"]) <= 1e-10
    assert rel(got[1], o.chem_potential[1]) <= 1e-8
    assert rel(got[2], 0.42 * V) <= 1e-12
