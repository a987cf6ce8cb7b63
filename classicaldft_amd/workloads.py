"""Seeded synthetic inputs shared by the benchmark, the warm-start pipeline and the tests.

No counterpart in the reference (its drivers read parameter files); these
generate the systems of the batched and warm-start workloads: per system a bulk
density U[0.3, 0.8] and beta V_ext = sum of 3 periodic Gaussian wells,
amplitude U[-3, 3], width U[0.3, 1.0], centre U[0, L)^3.
"""
import numpy as np


def splitmix64_uniform(n, seed):
    """n uniforms in [0, 1) from the SplitMix64 sequence started at `seed` (53-bit mantissas)."""
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def well_system(n, seed, dx=0.1, dtype=np.float64):
    """(bulk density, beta V_ext on the periodic n^3 grid) of system `seed`; minimum-image distances."""
    u = splitmix64_uniform(16, seed)
    L = n * dx
    x = np.arange(n) * dx
    v = np.zeros((n, n, n), dtype=dtype)
    for j in range(3):
        amp, w = -3 + 6 * u[1 + j], 0.3 + 0.7 * u[4 + j]
        c = u[7 + 3 * j: 10 + 3 * j] * L
        g = [np.exp(-np.minimum(np.abs(x - c[k]), L - np.abs(x - c[k])) ** 2 / (2 * w * w)) for k in range(3)]
        v += (amp * g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]).astype(dtype)
    return 0.3 + 0.5 * u[0], v


def well_systems(n, count, first_seed, dx=0.1, dtype=np.float64):
    """Systems first_seed .. first_seed + count - 1: arrays (bulk densities, beta V_ext)."""
    rho_b = np.empty(count)
    vexts = np.empty((count, n, n, n), dtype=dtype)
    for s in range(count):
        rho_b[s], vexts[s] = well_system(n, first_seed + s, dx, dtype)
    return rho_b, vexts
