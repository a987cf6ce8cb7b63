"""Pins of the oracle's Phi models (include/FMT.h:139-643, src/FMT.cpp:114-176, 616-709).

* Bulk closed forms: White Bear I and RSLT reduce to Carnahan-Starling,
  Rosenfeld to the Percus-Yevick compressibility route.
* Thermodynamic consistency: mu_ex = d f_ex / d rho for every model.
* dPhi against central differences over all 19 fundamental measures at random
  (non-uniform) points, tensor entries perturbed one at a time.
* The eta < 1e-12 series branches equal the exact Taylor expansions of the
  closed forms (coefficients from the series of 1/(1-x)^2 and ln(1-x)).
* Committed bulk table (tests/golden/bulk_table.json) guards regressions.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle

MODELS = [(oracle.ROSENFELD, 1.0, 0.0), (oracle.RSLT, 1.0, 0.0), (oracle.ESFMT, 1.3, -0.4), (oracle.WBI, 1.0, -1.0),
          (oracle.WBII, 1.0, -1.0)]


def cs_fex(rho, d=1.0):
    eta = np.pi * rho * d ** 3 / 6
    return rho * eta * (4 - 3 * eta) / (1 - eta) ** 2


def py_fex(rho, d=1.0):
    eta = np.pi * rho * d ** 3 / 6
    return rho * (-np.log(1 - eta) + 1.5 * eta * (2 - eta) / (1 - eta) ** 2)


@pytest.mark.parametrize("rho", [0.01, 0.2, 0.5, 0.8, 0.95])
def test_bulk_closed_forms(rho):
    for model in (oracle.WBI, oracle.RSLT):
        fex, _ = oracle.bulk(model, rho, 1.0)
        assert abs(fex / cs_fex(rho) - 1) < 1e-13
    fex, _ = oracle.bulk(oracle.ROSENFELD, rho, 1.0)
    assert abs(fex / py_fex(rho) - 1) < 1e-13


@pytest.mark.parametrize("model,A,B", MODELS)
def test_mu_is_derivative_of_f(model, A, B):
    for rho in (0.05, 0.3, 0.7):
        h = 1e-5 * rho
        fp, _ = oracle.bulk(model, rho + h, 1.3, A, B)
        fm, _ = oracle.bulk(model, rho - h, 1.3, A, B)
        _, mu = oracle.bulk(model, rho, 1.3, A, B)
        assert abs((fp - fm) / (2 * h) / mu - 1) < 1e-8


def random_measures(rng, eta=0.3, d=1.0):
    s2 = np.pi * d * d * eta * 6 / (np.pi * d ** 3) * (1 + 0.1 * rng.standard_normal())
    v2 = 0.2 * s2 * rng.standard_normal(3)
    T = s2 / 3 * np.eye(3) + 0.05 * s2 * rng.standard_normal((3, 3))
    T = 0.5 * (T + T.T)
    m = np.zeros(19)
    m[0] = eta
    m[1], m[2], m[3] = s2 / d ** 2, s2 / d, s2
    m[4:7] = v2 / d
    m[7:10] = v2
    m[10:] = T.reshape(-1)
    return m


@pytest.mark.parametrize("model,A,B", MODELS)
def test_dphi_central_differences(model, A, B):
    rng = np.random.default_rng(model + 7)
    for _ in range(3):
        m = random_measures(rng, eta=0.2 + 0.2 * rng.random())
        _, d = oracle.phi(model, m, A, B)
        for k in range(19):
            if model in (oracle.ROSENFELD, oracle.RSLT) and k >= 10:
                continue  # tensor measures do not enter these models
            h = 1e-6 * max(abs(m[k]), 1e-2)
            mp, mm = m.copy(), m.copy()
            mp[k] += h
            mm[k] -= h
            num = (oracle.phi(model, mp, A, B)[0] - oracle.phi(model, mm, A, B)[0]) / (2 * h)
            assert abs(num - d[k]) <= 1e-6 * max(abs(d).max(), 1e-12), (k, num, d[k])


def taylor_f3_rslt(n):
    # 1/(x(1-x)^2) + ln(1-x)/x^2 = sum_k ((k+2) - 1/(k+2)) x^k
    return [Fraction(k + 2) - Fraction(1, k + 2) for k in range(n)]


def taylor_f3_wbii(n):
    # -(1-3x+x^2)/(x(1-x)^2) - ln(1-x)/x^2
    # (1-3x+x^2)/(1-x)^2 = sum_k c_k x^k with c_k = (k+1) - 3k + (k-1) = -k for k >= 2, c_0 = 1, c_1 = -1
    c = [Fraction(1), Fraction(-1)] + [Fraction(-k) for k in range(2, n + 2)]
    # -(1/x) sum c_k x^k + sum_k x^k/(k+2) + 1/x  -> coefficient of x^k: -c_{k+1} + 1/(k+2)
    return [-c[k + 1] + Fraction(1, k + 2) for k in range(n)]


def taylor_f2_wbii(n):
    # 1/3 + (4/3)/(1-x) + (2/(3x)) ln(1-x) = sum_k (4/3 - 2/(3(k+1))) x^k  (+1/3 at k = 0)
    out = [Fraction(4, 3) - Fraction(2, 3 * (k + 1)) for k in range(n)]
    out[0] += Fraction(1, 3)
    return out


def test_series_branches_are_exact_taylor():
    x = 5e-13
    f = oracle.fcoef(oracle.WBI, x)
    t = taylor_f3_rslt(4)
    assert [float(v) for v in t] == [1.5, 8 / 3, 3.75, 4.8]
    assert abs(f[4] - sum(float(c) * x ** k for k, c in enumerate(t))) < 1e-15
    assert abs(f[5] - sum(k * float(c) * x ** (k - 1) for k, c in enumerate(t) if k)) < 1e-13
    t2 = taylor_f3_wbii(4)
    assert [float(v) for v in t2] == pytest.approx([1.5, 7 / 3, 3.25, 4.2], abs=1e-15)
    fw = oracle.fcoef(oracle.WBII, x)
    assert abs(fw[4] - sum(float(c) * x ** k for k, c in enumerate(t2))) < 1e-15
    t3 = taylor_f2_wbii(4)
    assert [float(v) for v in t3] == pytest.approx([1.0, 1.0, 10 / 9, 7 / 6], abs=1e-15)
    assert abs(fw[2] - sum(float(c) * x ** k for k, c in enumerate(t3))) < 1e-15


def test_closed_form_branches_away_from_zero():
    """At eta = 0.01 the closed forms agree with the Taylor series to its truncation error."""
    x = 0.01
    f = oracle.fcoef(oracle.WBI, x)
    t = taylor_f3_rslt(12)
    assert abs(f[4] - sum(float(c) * x ** k for k, c in enumerate(t))) < 1e-12
    fw = oracle.fcoef(oracle.WBII, x)
    t2 = taylor_f3_wbii(12)
    assert abs(fw[4] - sum(float(c) * x ** k for k, c in enumerate(t2))) < 1e-12


def test_bulk_table_regression():
    with open(os.path.join(os.path.dirname(__file__), "golden", "bulk_table.json")) as fh:
        table = json.load(fh)
    for model, rows in table.items():
        for eta, fex, mu in rows:
            rho = eta * 6 / np.pi
            f2, m2 = oracle.bulk(int(model), rho, 1.0)
            assert abs(f2 - fex) <= 1e-14 * max(1.0, abs(fex))
            assert abs(m2 - mu) <= 1e-14 * max(1.0, abs(mu))
