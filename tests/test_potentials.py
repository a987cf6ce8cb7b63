"""Pair potentials of the mean field (include/Potential1.h) on the host side of the
library, and the oracle's lattice kernels for every rule of include/Interaction.h.

Independent pins: closed forms of V and of the WCA / BH tails at sample radii,
scipy quadrature of Potential1::getHSD and getVDW_Parameter, numpy's
Gauss-Legendre rule, the continuum limit of the lattice kernel (its sum tends to
the integral of Watt over space, twice getVDW_Parameter, as dx -> 0), and the
kernel's lattice symmetries."""
import numpy as np
import pytest
from scipy import integrate

import oracle
import classicaldft_amd as cdft

KINDS = [(cdft.POT_LJ, {}), (cdft.POT_TWF, {"alpha": 50.0}), (cdft.POT_WHDF, {})]


def closed_form(kind, r, sigma=1.0, eps=1.0, rcut=2.5, alpha=50.0):
    """V(r) = vr(r) - shift and (r_min, r_0) of each potential, written from the formulas of the papers:
    LJ 4 eps ((s/r)^12 - (s/r)^6); ten Wolde-Frenkel 4 eps / alpha^2 (y^6 - alpha y^3), y = 1 / ((r/s)^2 - 1);
    Wang-Ramirez-Dobnikar-Frenkel eps' ((s/r)^2 - 1)((rc/r)^2 - 1)^2."""
    if kind == cdft.POT_LJ:
        vr = lambda x: 4 * eps * ((sigma / x) ** 12 - (sigma / x) ** 6)
        shift = vr(rcut)
        rmin = 2 ** (1 / 6) * sigma
        r0 = (0.5 * np.sqrt(1 + shift) + 0.5) ** (-1 / 6)
    elif kind == cdft.POT_TWF:
        def vr(x):
            with np.errstate(divide="ignore", invalid="ignore"):
                y = 1.0 / ((x / sigma) ** 2 - 1)
                return np.where(x < sigma, 1e50, 4 * eps / alpha ** 2 * (y ** 6 - alpha * y ** 3))
        shift = float(vr(np.array(rcut)))
        rmin = sigma * np.sqrt(1 + (2 / alpha) ** (1 / 3))
        r0 = np.sqrt(1 + (25 * np.sqrt(1 + shift) + 25) ** (-1 / 3))
    else:
        e2 = eps * 2 * (rcut / sigma) ** 2 * (2 * ((rcut / sigma) ** 2 - 1) / 3) ** -3
        vr = lambda x: e2 * ((sigma / x) ** 2 - 1) * ((rcut / x) ** 2 - 1) ** 2
        shift = 0.0
        rmin = rcut * ((1 + 2 * (rcut / sigma) ** 2) / 3) ** -0.5
        r0 = 1.0
    V = lambda x: vr(x) - shift
    return V, rmin, r0


@pytest.mark.parametrize("kind,kw", KINDS)
@pytest.mark.parametrize("bh", [False, True])
def test_attractive_tail_matches_closed_form(kind, kw, bh):
    V, rmin, r0 = closed_form(kind, None, **kw)
    r = np.linspace(0.3, 3.0, 541)
    got = cdft.potential_watt(r, kind=kind, bh=bh, **kw)
    if bh:
        want = np.where(r < r0, 0.0, V(r))
    else:
        want = np.where(r < rmin, V(rmin), V(r))
    want = np.where(r < 2.5, want, 0.0)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    # WCA continuation limit: zero inside r_att_min
    if not bh:
        w = cdft.potential_watt(r, kind=kind, r_att_min=0.9, **kw)
        assert np.all(w[r < 0.9] == 0.0) and np.array_equal(w[r >= 0.9], got[r >= 0.9])


@pytest.mark.parametrize("kind,kw", KINDS)
@pytest.mark.parametrize("bh", [False, True])
def test_hsd_and_vdw_match_quadrature(kind, kw, bh):
    kT = 1.1
    V, rmin, r0 = closed_form(kind, None, **kw)
    hc = 1.0 if kind == cdft.POT_TWF else 0.0
    rlim = r0 if bh else rmin
    V0 = (lambda r: V(r)) if bh else (lambda r: V(r) - V(rmin))
    d = hc + integrate.quad(lambda r: 1 - np.exp(-V0(r) / kT) if r > 0 else 1.0, hc, rlim, epsabs=1e-13,
                            epsrel=1e-13, limit=200)[0]
    assert abs(cdft.potential_hsd(kT, kind=kind, bh=bh, **kw) - d) <= 1e-10
    tail = (lambda r: r * r * V(r))
    if bh:
        a = integrate.quad(tail, r0, 2.5, epsabs=1e-14, epsrel=1e-13, limit=200)[0]
    else:
        a = integrate.quad(lambda r: r * r * V(rmin), 0.0, rmin, epsabs=1e-14)[0] + \
            integrate.quad(tail, rmin, 2.5, epsabs=1e-14, epsrel=1e-13, limit=200)[0]
    assert abs(cdft.potential_vdw(kT, kind=kind, bh=bh, **kw) - 2 * np.pi / kT * a) <= 1e-10 * abs(a)
    assert cdft.lj_hsd(kT) == cdft.potential_hsd(kT)


def test_gauss_legendre_rule():
    for n in (1, 2, 3, 5, 8):
        x, w = oracle.gauss_legendre01(n)
        for k in range(2 * n):  # exact for polynomials of degree 2n - 1 on [0, 1]
            assert abs(np.sum(w * x ** k) - 1.0 / (k + 1)) <= 1e-14


@pytest.mark.parametrize("scheme", range(7))
def test_kernel_continuum_limit_and_symmetry(scheme):
    """Sum_R w_att(R) -> int Watt d^3r / kT = 2 getVDW_Parameter as dx -> 0; w_att(R) = w_att(-R) and is invariant
    under axis permutations on a cubic lattice."""
    kT = 1.0
    a_cont = 2 * cdft.potential_vdw(kT)
    errs = []
    for dx in (0.2, 0.1, 0.05):
        n = int(round(6.0 / dx))
        n += n % 2
        w, a = oracle.mf_kernel2((n, n, n), dx, scheme=scheme, ngauss=3, kT=kT)
        assert abs(np.sum(w) - a) <= 1e-12 * abs(a)
        errs.append(abs(a - a_cont) / abs(a_cont))
        if dx == 0.2:
            assert np.allclose(w, np.roll(np.flip(w, (0, 1, 2)), 1, (0, 1, 2)), rtol=0, atol=1e-14 * np.abs(w).max())
            assert np.allclose(w, w.transpose(1, 2, 0), rtol=0, atol=1e-14 * np.abs(w).max())
    assert errs[-1] < 2e-3 and errs[-1] < errs[0], errs
