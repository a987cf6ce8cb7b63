"""Mean-field kernels of every potential, split and lattice rule generated on the
device against the oracle's restatement of generateWeightsXYZSym / generateWeight
(src/Interaction.cpp:178-307, 443-534; include/Interaction.h:293-431), and a full
DFT evaluation with a non-default potential and rule against the oracle DFT."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SHAPE, DX = (24, 24, 24), 0.2
POTS = [(0, {}), (1, {"alpha": 50.0}), (1, {"alpha": 30.0}), (2, {})]


def maxrel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("scheme", range(7))
@pytest.mark.parametrize("kind,kw", POTS)
@pytest.mark.parametrize("bh", [False, True])
def test_device_kernel_matches_oracle(cdft, scheme, kind, kw, bh):
    kT = 0.9
    ng = 3 if scheme >= 5 else 0
    w0, a0 = oracle.mf_kernel2(SHAPE, DX, kind=kind, bh=bh, scheme=scheme, ngauss=ng, kT=kT, **kw)
    s = cdft.System(SHAPE, DX, 1.0, mean_field=True, kT=kT, potential=kind, bh=bh, mf_scheme=scheme, ngauss=ng,
                    alpha=kw.get("alpha", 0.0))
    w1 = s.device_mf_kernel()
    assert maxrel(w1, w0) <= 1e-13, maxrel(w1, w0)
    assert abs(s.a_vdw - a0) <= 1e-12 * abs(a0)
    s.close()


def test_wca_continuation_limit(cdft):
    w0, a0 = oracle.mf_kernel2(SHAPE, DX, r_att_min=0.95, scheme=cdft.MF_LF, kT=1.0)
    s = cdft.System(SHAPE, DX, 1.0, mean_field=True, kT=1.0, r_att_min=0.95, mf_scheme=cdft.MF_LF)
    assert maxrel(s.device_mf_kernel(), w0) <= 1e-13


def test_full_dft_with_twf_gauss_e(cdft):
    """Ideal + White Bear I + tWF (BH split, Gauss_E kernel, 2 points) + V_ext - mu against the oracle DFT."""
    kT, mu = 1.2, -0.7
    rng = np.random.default_rng(9)
    rho = 0.5 * (1 + 0.1 * (2 * rng.random(SHAPE) - 1))
    z = np.arange(SHAPE[2]) * DX
    vext = np.broadcast_to(np.exp(-z / 0.3), SHAPE).copy()
    w, a = oracle.mf_kernel2(SHAPE, DX, kind=cdft.POT_TWF, alpha=50.0, bh=True, scheme=cdft.MF_GAUSS_E, ngauss=2,
                             kT=kT)
    wf = oracle.fmt_weights(SHAPE, DX, 1.0)
    o = oracle.DFT(SHAPE, DX, 1.0, model=oracle.WBI, mf=True, kT=kT, mu=mu, vext=vext)
    o.set_mean_field(w, a)
    p0, d0 = o.eval(rho)
    s = cdft.System(SHAPE, DX, 1.0, model=cdft.WHITEBEAR_I, weights=wf, ideal=True, mu=mu, vext=vext,
                    mean_field=True, kT=kT, potential=cdft.POT_TWF, alpha=50.0, bh=True, mf_scheme=cdft.MF_GAUSS_E,
                    ngauss=2)
    p1, d1 = s.eval(rho)
    for k in range(5):
        assert abs(p1[k] - p0[k]) <= 1e-11 * abs(p0[0])
    assert maxrel(d1, d0) <= 1e-11
    # bulk mu uses the kernel's a_vdw (src/Interaction.cpp:61-72)
    _, muex = oracle.bulk(oracle.WBI, 0.5, 1.0)
    assert abs(s.bulk(0.5)[1] - (np.log(0.5) + muex + a * 0.5)) <= 1e-12 * abs(s.bulk(0.5)[1])
