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
