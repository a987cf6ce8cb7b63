"""Density files in the reference's format (no GPU work).

``Density::writeDensity`` / ``readDensity`` (src/Density.cpp:227-244) store a
density through ``DFT_Vec::save`` / ``load`` (additionalCode/DFT_Eigen.cpp:71-85):
text, the number of values on the first line, then the values in lattice order
(z fastest) separated by one blank, each printed by a default-formatted C++
stream (6 significant digits, ``%g``), then a newline.  FIRE writes such files
as ``snapshot_s<species>.dat`` every 10 steps (src/Minimizer.cpp:259-268).
"""
import numpy as np


def write_density(path, rho, precision=6):
    """Write rho in the reference's text format (precision: significant digits; 6 is what the reference
    writes, 17 makes the file round-trip exactly)."""
    flat = np.ascontiguousarray(rho, dtype=np.float64).ravel()
    fmt = "%." + str(int(precision)) + "g"
    with open(path, "w", newline="\n") as f:
        f.write(f"{flat.size}\n")
        for b in range(0, flat.size, 1 << 16):
            chunk = flat[b:b + (1 << 16)]
            f.write("".join(fmt % x + " " for x in chunk.tolist()))
        f.write("\n")


def read_density(path, shape=None):
    """Read a reference-format density file; with shape, check the count and reshape."""
    with open(path, "r") as f:
        head = f.readline()
        n = int(head.split()[0])
        vals = np.array(f.read().split(), dtype=np.float64)
    if vals.size != n:
        raise ValueError(f"{path}: header says {n} values, found {vals.size}")
    if shape is not None:
        if int(np.prod(shape)) != n:
            raise ValueError(f"{path}: {n} values do not fill a {tuple(shape)} lattice")
        vals = vals.reshape(shape)
    return vals
