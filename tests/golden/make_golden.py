"""Writes the small golden fixtures under tests/golden/.

* fft_vectors.npz  - numpy.fft.rfftn of seeded real arrays (pins the FFT conventions)
* config1_wb_32.npz - config 1 (32^3, dx = 0.1, d = 1, White Bear I, rho = 0.6 (1 + 0.05 U[-1,1]),
                      seed 0x5EED0001): rho seed, the oracle's F and a dF checksum, for both
                      the reference-faithful and the accurate weight sets.
* bulk_table.json  - f_ex and mu_ex per model on a grid of packing fractions (oracle).

The FMT numbers come from the repository's oracle (a restatement of the
reference; the reference itself cannot be built here), so they pin the oracle
against regressions, not against the reference; the closed forms in
tests/test_oracle_phi.py pin it against the mathematics.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def main():
    rng = np.random.default_rng(20240601)
    fft = {}
    for shape in [(4, 4, 4), (8, 8, 8), (6, 12, 18), (16, 16, 16)]:
        tag = "x".join(map(str, shape))
        x = rng.standard_normal(shape)
        fft["in_" + tag] = x
        fft["r2c_" + tag] = np.fft.rfftn(x)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **fft)

    shape, dx, hsd = (32, 32, 32), 0.1, 1.0
    u = oracle.splitmix64_uniform(32 ** 3, 0x5EED0001).reshape(shape)
    rho = 0.6 * (1 + 0.05 * (2 * u - 1))
    out = {"rho_head": rho.reshape(-1)[:16]}
    for name, mode in [("faithful", oracle.WEIGHTS_FAITHFUL), ("accurate", oracle.WEIGHTS_ACCURATE)]:
        w = oracle.fmt_weights(shape, dx, hsd, mode=mode)
        F, dF = oracle.fmt_eval(oracle.WBI, rho, dx, hsd, weights=w)
        out["F_" + name] = np.array(F)
        out["dF_sum_" + name] = np.array(dF.sum())
        out["dF_corner_" + name] = dF[:2, :2, :4].copy()
        out["w_sums_" + name] = w.sum(axis=(1, 2, 3))
    np.savez_compressed(os.path.join(HERE, "config1_wb_32.npz"), **out)

    table = {}
    for model in range(5):
        rows = []
        for eta in [1e-3, 0.05, 0.1, 0.2, 0.3, 0.4, 0.45, 0.5]:
            rho = eta * 6 / np.pi
            fex, mu = oracle.bulk(model, rho, 1.0)
            rows.append([eta, fex, mu])
        table[str(model)] = rows
    with open(os.path.join(HERE, "bulk_table.json"), "w") as f:
        json.dump(table, f, indent=1)


if __name__ == "__main__":
    main()
