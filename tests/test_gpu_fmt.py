"""GPU parity of the FMT / DFT hot path against the CPU oracle (through the C ABI).

Tolerances: the FP64 path must match the oracle to 1e-11 relative in F and in
max-norm dF.  The weights are an input of the
comparison: the reference's double-precision weight generator carries
roundoff that alone moves F by ~1e-10 at dx = 0.1 (see
tests/test_oracle_weights.py), so the FP64 parity cases hand both sides the
same weights, and separate cases check the device generator against the
oracle's accurate weights (well-conditioned integrals in IEEE quad), and the
reference-expression mode against the oracle's reference-faithful weights.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-11


def seeded_density(shape, rho_b, seed, amp=0.05):
    u = oracle.splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    return rho_b * (1.0 + amp * (2.0 * u - 1.0))


def rel(a, b):
    return abs(a - b) / abs(b)


def maxrel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 12, 18), (32, 32, 32), (48, 48, 48), (24, 16, 96), (64, 64, 64),
                                   (256, 6, 8), (4, 256, 12), (128, 96, 4), (96, 128, 6), (12, 24, 12), (6, 4, 16),
                                   (4, 6, 512), (6, 4, 256), (4, 4, 192), (4, 8, 128), (3, 6, 36)])
def test_fft_matches_numpy(cdft, shape):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape)
    ref = np.fft.rfftn(x)
    got = cdft.fft_r2c(x)
    assert maxrel(got, ref) < 1e-13
    back = cdft.fft_c2r(ref, shape[2])
    assert maxrel(back, x * x.size) < 1e-13
    gotm = cdft.fft_r2c(x, precision=cdft.MIXED)
    assert maxrel(gotm, ref) < 2e-6


def test_config1_fp64_parity_shared_weights(cdft):
    """Config 1: 32^3, dx = 0.1, d = 1, White Bear I, rho = 0.6 (1 + 0.05 U[-1,1])."""
    shape, dx, hsd = (32, 32, 32), 0.1, 1.0
    rho = seeded_density(shape, 0.6, 0x5EED0001)
    w = oracle.fmt_weights(shape, dx, hsd)
    F0, dF0 = oracle.fmt_eval(oracle.WBI, rho, dx, hsd, weights=w)
    sysm = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, weights=w)
    F1, dF1 = sysm.fmt_eval(rho)
    print("config1 F", F0, F1, rel(F1, F0), "dF", maxrel(dF1, dF0))
    assert rel(F1, F0) <= TOL
    assert maxrel(dF1, dF0) <= TOL


@pytest.mark.parametrize("model,A,B", [(0, 1.0, 0.0), (1, 1.0, 0.0), (2, 1.3, -0.4), (3, 1.0, -1.0), (4, 1.0, -1.0)])
def test_all_models_fp64_parity(cdft, model, A, B):
    shape, dx, hsd = (16, 16, 16), 0.1, 1.0
    rho = seeded_density(shape, 0.5, 0x5EED0100 + model, amp=0.2)
    w = oracle.fmt_weights(shape, dx, hsd)
    F0, dF0 = oracle.fmt_eval(model, rho, dx, hsd, weights=w, A=A, B=B)
    s = cdft.System(shape, dx, hsd, model=model, weights=w, A=A, B=B)
    F1, dF1 = s.fmt_eval(rho)
    assert rel(F1, F0) <= TOL
    assert maxrel(dF1, dF0) <= TOL


@pytest.mark.parametrize("shape", [(128, 24, 16), (12, 256, 24), (96, 48, 12), (16, 12, 256), (12, 16, 192),
                                   (12, 12, 512)])
def test_long_axes_fp64_parity(cdft, shape):
    """Transforms of lengths 512 (z), 256, 192 (z), 128, 96, 48, 24, 12 inside the full F + dF pipeline."""
    dx, hsd = 0.1, 1.0
    rho = seeded_density(shape, 0.55, 0x5EED0300 + shape[0], amp=0.1)
    w = oracle.fmt_weights(shape, dx, hsd)
    F0, dF0 = oracle.fmt_eval(oracle.WBI, rho, dx, hsd, weights=w)
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, weights=w)
    F1, dF1 = s.fmt_eval(rho)
    print(shape, "F", rel(F1, F0), "dF", maxrel(dF1, dF0))
    assert rel(F1, F0) <= TOL
    assert maxrel(dF1, dF0) <= TOL


@pytest.mark.parametrize("shape,dx", [((32, 32, 32), 0.1), ((64, 64, 64), 0.04), ((48, 36, 64), 0.05)])
def test_device_weights_match_accurate_oracle(cdft, shape, dx):
    """The double-double device generator against the quad oracle, cell by cell.  In plain double the
    cancelling corner sums leave 4e-11 dV (dx = 0.1) .. 1e-8 dV (dx = 0.04); here every cell is the correctly
    rounded double up to the last bit."""
    hsd = 1.0
    s = cdft.System(shape, dx, hsd)
    wg = s.device_weights()
    wa = oracle.fmt_weights(shape, dx, hsd, mode=oracle.WEIGHTS_ACCURATE)
    dV = dx ** 3
    err = np.abs(wg - wa).max() / dV
    print(shape, dx, "max |w_gpu - w_accurate| / dV =", err)
    assert err <= 1e-15
    # sum rules of the trilinear-hat partition of unity
    assert abs(wg[0].sum() / (np.pi * hsd ** 3 / 6) - 1) < 1e-13
    assert abs(wg[1].sum() / (np.pi * hsd ** 2) - 1) < 1e-13


@pytest.mark.parametrize("shape,dx", [((32, 32, 32), 0.1), ((48, 48, 48), 0.04)])
def test_reference_weights_mode_matches_faithful_oracle(cdft, shape, dx):
    """CDFT_WEIGHTS_REFERENCE: the reference's own double expressions (asin forms, R += 1e-15), evaluated on the
    host operation for operation, equal the oracle's reference-faithful transcription bit for bit, noise
    included; with them the drop-in reproduces the oracle's reference-faithful F and dF."""
    hsd = 1.0
    s = cdft.System(shape, dx, hsd, weights_mode=cdft.WEIGHTS_REFERENCE)
    wr = s.device_weights()
    wf = oracle.fmt_weights(shape, dx, hsd, mode=oracle.WEIGHTS_FAITHFUL)
    assert np.array_equal(wr, wf)
    wa = oracle.fmt_weights(shape, dx, hsd, mode=oracle.WEIGHTS_ACCURATE)
    assert np.abs(wr - wa).max() > 1e-12 * dx ** 3  # the reference's rounding noise is there
    rho = seeded_density(shape, 0.6, 0x5EED0001)
    F0, dF0 = oracle.fmt_eval(oracle.WBI, rho, dx, hsd, weights=wf)
    F1, dF1 = s.fmt_eval(rho)
    print(shape, "reference-mode F", rel(F1, F0), "dF", maxrel(dF1, dF0))
    assert rel(F1, F0) <= TOL
    assert maxrel(dF1, dF0) <= TOL


def test_metric_grid_parity_shared_faithful_weights(cdft):
    """The metric workload: 256^3, dx = 0.04, White Bear I, rho = 0.6 (1 + 0.05 U[-1,1]), seed 0x5EED0256, both
    sides on the reference-faithful weights."""
    shape, dx, hsd = (256, 256, 256), 0.04, 1.0
    rho = seeded_density(shape, 0.6, 0x5EED0256)
    w = oracle.fmt_weights(shape, dx, hsd, mode=oracle.WEIGHTS_FAITHFUL)
    F0, dF0 = oracle.fmt_eval(oracle.WBI, rho, dx, hsd, weights=w)
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, weights=w)
    del w
    F1, dF1 = s.fmt_eval(rho)
    s.close()
    print("256^3 shared weights F", F0, F1, rel(F1, F0), "dF", maxrel(dF1, dF0))
    assert rel(F1, F0) <= TOL
    assert maxrel(dF1, dF0) <= TOL


def test_metric_grid_parity_own_weights(cdft):
    """The metric workload with the product's own device-generated weights against the oracle on its accurate
    (quad) weights."""
    shape, dx, hsd = (256, 256, 256), 0.04, 1.0
    rho = seeded_density(shape, 0.6, 0x5EED0256)
    wa = oracle.fmt_weights(shape, dx, hsd, mode=oracle.WEIGHTS_ACCURATE)
    F0, dF0 = oracle.fmt_eval(oracle.WBI, rho, dx, hsd, weights=wa)
    del wa
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I)
    F1, dF1 = s.fmt_eval(rho)
    s.close()
    print("256^3 own weights F", F0, F1, rel(F1, F0), "dF", maxrel(dF1, dF0))
    assert rel(F1, F0) <= TOL
    assert maxrel(dF1, dF0) <= TOL


def test_config1_fp64_parity_own_weights(cdft):
    """End to end: device-generated weights vs the oracle on its accurate weights."""
    shape, dx, hsd = (32, 32, 32), 0.1, 1.0
    rho = seeded_density(shape, 0.6, 0x5EED0001)
    wa = oracle.fmt_weights(shape, dx, hsd, mode=oracle.WEIGHTS_ACCURATE)
    F0, dF0 = oracle.fmt_eval(oracle.WBI, rho, dx, hsd, weights=wa)
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I)
    F1, dF1 = s.fmt_eval(rho)
    print("own weights F", rel(F1, F0), "dF", maxrel(dF1, dF0))
    assert rel(F1, F0) <= TOL
    assert maxrel(dF1, dF0) <= TOL


def test_uniform_density_closed_form(cdft):
    """Uniform rho: F/V = f_ex(rho) (Carnahan-Starling for WB I) and dF = dV mu_ex."""
    shape, dx, hsd, rb = (32, 32, 32), 0.1, 1.0, 0.7
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I)
    F, dF = s.fmt_eval(np.full(shape, rb))
    eta = np.pi * rb / 6
    cs = rb * eta * (4 - 3 * eta) / (1 - eta) ** 2
    V = np.prod(shape) * dx ** 3
    fex, muex = oracle.bulk(oracle.WBI, rb, hsd)
    assert abs(F / V - cs) / cs < 1e-11
    assert abs(fex - cs) / cs < 1e-13
    assert np.abs(dF / dx ** 3 - muex).max() / muex < 1e-11


def test_full_dft_parity(cdft):
    """Ideal + WB I + LJ/QF mean field + V_ext - mu against the oracle DFT."""
    shape, dx, hsd = (32, 32, 32), 0.1, 1.0
    rho = seeded_density(shape, 0.5, 0x5EED0200, amp=0.1)
    z = np.arange(shape[2]) * dx
    vext = np.broadcast_to(np.exp(-z / 0.2) + np.exp(-(shape[2] * dx - z) / 0.2), shape).copy()
    kT, mu = 0.8, -1.3
    w = oracle.fmt_weights(shape, dx, hsd)
    o = oracle.DFT(shape, dx, hsd, model=oracle.WBI, mf=True, kT=kT, mu=mu, vext=vext)
    p0, dF0 = o.eval(rho)
    wk, a0 = oracle.mf_kernel(shape, dx, kT=kT)
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, weights=w, ideal=True, mu=mu, vext=vext, mean_field=True,
                    kT=kT)
    assert maxrel(s.device_mf_kernel(), wk) < 1e-13
    assert rel(s.a_vdw, a0) < 1e-13
    assert rel(s.a_vdw, o.a_vdw) < 1e-13
    p1, dF1 = s.eval(rho)
    print("parts oracle", p0, "gpu", p1)
    for k in range(5):
        assert abs(p1[k] - p0[k]) <= TOL * abs(p0[0]) + 1e-300
    assert maxrel(dF1, dF0) <= TOL


def test_eta_too_large_is_reported(cdft):
    shape, dx = (16, 16, 16), 0.1
    rho = np.full(shape, 0.3)
    rho[3, 4, 5] = 2.0 / dx ** 3 * 10
    s = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I)
    with pytest.raises(cdft.EtaTooLarge):
        s.fmt_eval(rho)
    with pytest.raises(oracle.EtaTooLarge):
        oracle.fmt_eval(oracle.WBI, rho, dx, 1.0)
    # the system stays usable afterwards
    F, _ = s.fmt_eval(np.full(shape, 0.3))
    assert np.isfinite(F)


def test_accumulate_flag(cdft):
    shape, dx = (16, 16, 16), 0.1
    rho = seeded_density(shape, 0.4, 7)
    s = cdft.System(shape, dx, 1.0)
    _, d1 = s.fmt_eval(rho)
    base = np.full(shape, 3.0)
    _, d2 = s.fmt_eval(rho, dF=base.copy(), accumulate=True)
    assert np.abs(d2 - 3.0 - d1).max() < 1e-14 * np.abs(d1).max() + 1e-15


# Mixed precision against the FP64 oracle: |F error| <= 1e-7 relative; max-norm dF within 1e-6.
MIXED_F_TOL, MIXED_DF_TOL = 1e-7, 1e-6


@pytest.mark.parametrize("n,dx,seed", [(32, 0.1, 0x5EED0001), (256, 0.04, 0x5EED0256)])
def test_mixed_precision_error_vs_oracle(cdft, n, dx, seed):
    """Config 1 (32^3) and the metric workload (256^3) in the mixed path (FP32 transforms and convolutions, FP64
    Phi, the weighted densities' and dPhi's means in FP64) against the FP64 oracle on accurate weights."""
    shape, hsd = (n, n, n), 1.0
    rho = seeded_density(shape, 0.6, seed)
    oracle.set_threads(16, 16)
    try:
        wa = oracle.fmt_weights(shape, dx, hsd, mode=oracle.WEIGHTS_ACCURATE)
        F0, dF0 = oracle.fmt_eval(oracle.WBI, rho, dx, hsd, weights=wa)
    finally:
        oracle.set_threads(1, 1)
    del wa
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, precision=cdft.MIXED)
    F1, dF1 = s.fmt_eval(rho)
    s.close()
    print(f"{n}^3 mixed vs oracle: F rel {rel(F1, F0):.3e}, dF maxrel {maxrel(dF1, dF0):.3e}")
    assert rel(F1, F0) <= MIXED_F_TOL
    assert maxrel(dF1, dF0) <= MIXED_DF_TOL


def test_mixed_full_dft_vs_oracle(cdft):
    """The whole DFT (ideal gas, White Bear I, LJ mean field, V_ext, mu) in the mixed path against the oracle: the
    mean-field convolution's mean and the FMT force's mean travel in FP64 too."""
    shape, dx, hsd, kT = (32, 32, 32), 0.1, cdft.lj_hsd(0.8), 0.8
    rho = seeded_density(shape, 0.6, 0x5EED0007, amp=0.2)
    rng = np.random.default_rng(7)
    vext = rng.uniform(0.0, 1.0, shape)
    o = oracle.DFT(shape, dx, hsd, mf=True, kT=kT, mu=-1.5, vext=vext, accurate_weights=True)
    p0, dF0 = o.eval(rho)
    s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=kT, mu=-1.5, vext=vext,
                    precision=cdft.MIXED)
    p1, dF1 = s.eval(rho)
    s.close()
    print("full DFT mixed vs oracle: F rel", rel(p1[0], p0[0]), "dF", maxrel(dF1, dF0))
    assert rel(p1[0], p0[0]) <= MIXED_F_TOL
    assert maxrel(dF1, dF0) <= MIXED_DF_TOL


@pytest.mark.parametrize("precision", [0, 1])
def test_register_fft_matches_lds_fft(cdft, monkeypatch, precision):
    """The register two-pass FFT kernels and the LDS Stockham kernels give the same F and dF."""
    shape, dx, hsd = (64, 48, 32), 0.1, 1.0
    rho = seeded_density(shape, 0.6, 0x5EED0400)
    F1, dF1 = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, precision=precision).fmt_eval(rho)
    monkeypatch.setenv("CDFT_LDS_FFT", "1")
    F0, dF0 = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, precision=precision).fmt_eval(rho)
    tol = 1e-13 if precision == 0 else 2e-6
    print("reg vs lds", rel(F1, F0), maxrel(dF1, dF0))
    assert rel(F1, F0) <= tol
    assert maxrel(dF1, dF0) <= tol


def test_x_slab_section_matches_whole_array_passes(cdft, monkeypatch):
    """The y / z / y section run x-slab by x-slab (a small cache budget forces several slabs) gives the same
    F and dF as three whole-array passes, FP64 and mixed, single systems and batches."""
    shape, dx = (32, 24, 48), 0.1
    rho = seeded_density(shape, 0.5, 0x5EED0300, amp=0.1)
    kw = dict(model=cdft.WHITEBEAR_I, ideal=True, mean_field=True, kT=1.0, mu=0.2)
    for prec in (cdft.FP64, cdft.MIXED):
        monkeypatch.setenv("CDFT_SLAB_MB", "0")
        p0, d0 = cdft.System(shape, dx, 1.0, precision=prec, **kw).eval(rho)
        monkeypatch.setenv("CDFT_SLAB_MB", "1")  # 1 MiB: a few x-planes per slab
        p1, d1 = cdft.System(shape, dx, 1.0, precision=prec, **kw).eval(rho)
        assert np.max(np.abs(p1 - p0)) <= 1e-14 * abs(p0[0])
        assert maxrel(d1, d0) <= 1e-14
        b = cdft.System(shape, dx, 1.0, precision=prec, mus=[0.2, 0.3], model=cdft.WHITEBEAR_I, ideal=True,
                        mean_field=True, kT=1.0)
        pb, db = b.eval(np.array([rho, rho]))
        assert np.max(np.abs(pb[0] - p0)) <= 1e-14 * abs(p0[0])
        assert maxrel(db[0], d0) <= 1e-14


def test_weights_without_axis_parity_keep_the_full_spectra(cdft):
    """Caller weights for eta and s that stay even under r -> -r (so their spectra stay real, as the device stores
    them) but lose the evenness in x alone: the folded storage of the weight spectra must not be used (a fold would
    be off by the 1e-3 asymmetry); the result matches the oracle on those weights."""
    shape, dx, hsd = (16, 16, 16), 0.1, 1.0
    w = oracle.fmt_weights(shape, dx, hsd)
    u = oracle.splitmix64_uniform(int(np.prod(shape)), 0x5EED0777).reshape(shape)
    u = 0.5 * (u + np.roll(u[::-1, ::-1, ::-1], 1, axis=(0, 1, 2)))  # u(-r) = u(r)
    for a in (0, 1):
        w[a] = w[a] * (1 + 1e-3 * (2 * u - 1))
    rho = seeded_density(shape, 0.5, 0x5EED0778)
    F0, dF0 = oracle.fmt_eval(oracle.WBI, rho, dx, hsd, weights=w)
    for prec, tol in ((cdft.FP64, TOL), (cdft.MIXED, 1e-6)):
        s = cdft.System(shape, dx, hsd, model=cdft.WHITEBEAR_I, weights=w, precision=prec)
        F1, dF1 = s.fmt_eval(rho)
        assert rel(F1, F0) <= tol and maxrel(dF1, dF0) <= 10 * tol


@pytest.mark.parametrize("model", [3, 4, 2])
def test_trace_reduced_weight_set_matches_full(cdft, model, monkeypatch):
    """Weights satisfying w_Txx + w_Tyy + w_Tzz = w_s run with 10 stored arrays (n_Tzz rebuilt per site, its
    derivative folded back); the same system with all 11 arrays (CDFT_NO_TRACE=1) gives the same F, dF and
    Hessian product to rounding."""
    shape, dx, hsd = (48, 48, 64), 0.05, 1.0
    rho = seeded_density(shape, 0.55, 0x5EED0400 + model, amp=0.2)
    v = np.random.default_rng(model).standard_normal(shape)
    kw = dict(model=model, A=1.3, B=-0.4, ideal=True, mean_field=True, kT=1.2)
    res = []
    for flag in ("0", "1"):
        monkeypatch.setenv("CDFT_NO_TRACE", flag)
        s = cdft.System(shape, dx, hsd, **kw)
        parts, dF = s.eval(rho)
        h = s.hessian_dot(rho, v)
        res.append((parts[0], dF, h))
        s.close()
    (F0, dF0, h0), (F1, dF1, h1) = res
    print("trace vs full: F", rel(F0, F1), "dF", maxrel(dF0, dF1), "H v", maxrel(h0, h1))
    assert rel(F0, F1) <= 1e-13
    assert maxrel(dF0, dF1) <= 1e-13
    assert maxrel(h0, h1) <= 1e-12
