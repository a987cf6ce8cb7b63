"""Host side of the warm start: the committed trained parameters and the seeded workloads."""
import os

import numpy as np

from classicaldft_amd import workloads
from oracle import unet

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "classicaldft_amd", "data",
                       "warmstart_64.npz")


def test_committed_parameters():
    with np.load(WEIGHTS, allow_pickle=False) as z:
        p = z["params"]
        assert int(z["D"]) == 64
        assert float(z["kT"]) == 1.5
        assert int(z["train_systems"]) >= 4000
    assert p.dtype == np.float32 and p.size == unet.offsets()[3]
    assert np.all(np.isfinite(p))


def test_well_systems_match_direct_sum():
    """The separable evaluation of the three wells equals the direct minimum-image sum."""
    n, dx = 24, 0.1
    rho_b, v = workloads.well_system(n, 0x5EED5000 + 7, dx)
    u = workloads.splitmix64_uniform(16, 0x5EED5000 + 7)
    L = n * dx
    g = np.meshgrid(*[np.arange(n) * dx] * 3, indexing="ij")
    ref = np.zeros((n, n, n))
    for j in range(3):
        c = u[7 + 3 * j: 10 + 3 * j] * L
        d2 = sum(np.minimum(np.abs(g[k] - c[k]), L - np.abs(g[k] - c[k])) ** 2 for k in range(3))
        ref += (-3 + 6 * u[1 + j]) * np.exp(-d2 / (2 * (0.3 + 0.7 * u[4 + j]) ** 2))
    assert rho_b == 0.3 + 0.5 * u[0]
    assert np.max(np.abs(v - ref)) <= 1e-13


def test_splitmix64_known_values():
    """SplitMix64 from seed 0: first outputs 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4 (published test values)."""
    u = workloads.splitmix64_uniform(2, 0)
    assert u[0] == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert u[1] == (0x6E789E6AA1B965F4 >> 11) * 2.0 ** -53
