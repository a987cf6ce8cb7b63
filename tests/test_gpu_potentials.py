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


@pytest.mark.parametrize("kind,bh", [(0, False), (1, True), (2, False)])
def test_points_kernel_matches_oracle(cdft, kind, bh):
    """Interaction::generateWeights (src/Interaction.cpp:672-860) on the GPU against the serial restatement: the
    cells agree to rounding (atomic accumulation order), the total exactly up to rounding."""
    shape, dx = (24, 18, 16), 0.15
    i = np.arange(40) + 0.5
    z = 1 - 2 * i / 40
    phi = np.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    pts = np.stack([r * np.cos(phi), r * np.sin(phi), z, np.full(40, 4 * np.pi / 40)], axis=1)
    rc = 2.0
    w0 = oracle.mf_kernel_points(shape, dx, pts, kind=kind, rcut=rc, alpha=50.0, bh=bh, nradial=512)
    w1 = cdft.mf_kernel_points(shape, dx, pts, potential=kind, rcut=rc, alpha=50.0, bh=bh)
    m = np.abs(w0).max()
    assert np.max(np.abs(w1 - w0)) <= 1e-13 * m
    assert abs(w1.sum() - w0.sum()) <= 1e-12 * np.abs(w0).sum()
    # a full DFT with the kernel (divided by kT) as the mean field
    kT = 1.3
    rho = 0.4 * (1 + 0.1 * (2 * oracle.splitmix64_uniform(int(np.prod(shape)), 5).reshape(shape) - 1))
    s = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I, weights=oracle.fmt_weights(shape, dx, 1.0), ideal=True,
                    mean_field=True, kT=kT, mf_kernel=w1 / kT)
    o = oracle.DFT(shape, dx, 1.0, model=oracle.WBI, kT=kT)
    o.set_mean_field(w1 / kT, w1.sum() / kT)
    p0, d0 = o.eval(rho)
    p1, d1 = s.eval(rho)
    assert abs(p1[0] - p0[0]) <= 1e-11 * abs(p0[0])
    assert np.max(np.abs(d1 - d0)) <= 1e-11 * np.abs(d0).max()
    assert abs(s.a_vdw - w1.sum() / kT) <= 1e-12 * abs(w1.sum() / kT)


def test_scaled_mean_field(cdft):
    """scale_interaction's factor (desc mf_scale) on the kernel spectrum and a_vdw, against the oracle's DFT with the
    kernel and a_vdw scaled."""
    shape, dx, kT, f = (16, 12, 24), 0.1, 1.0, 0.7
    rho = 0.5 * (1 + 0.1 * (2 * oracle.splitmix64_uniform(int(np.prod(shape)), 3).reshape(shape) - 1))
    w, a = oracle.mf_kernel(shape, dx, kT=kT)
    o = oracle.DFT(shape, dx, 1.0, model=oracle.WBI, kT=kT, mu=0.2)
    o.set_mean_field(f * w, f * a)
    s = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I, weights=oracle.fmt_weights(shape, dx, 1.0), ideal=True,
                    mu=0.2, mean_field=True, kT=kT, mf_scale=f)
    assert abs(s.a_vdw - f * a) <= 1e-12 * abs(f * a)
    assert np.max(np.abs(s.device_mf_kernel() - f * w)) <= 1e-13 * np.abs(w).max()
    p0, d0 = o.eval(rho)
    p1, d1 = s.eval(rho)
    assert abs(p1[0] - p0[0]) <= 1e-11 * abs(p0[0]) and abs(p1[3] - p0[3]) <= 1e-11 * abs(p0[0])
    assert np.max(np.abs(d1 - d0)) <= 1e-11 * np.abs(d0).max()
