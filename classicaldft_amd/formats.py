"""Density files and mean-field weight caches in the reference's formats (no GPU work).

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


# ---------------------------------------------------------------------------
# Points-file mean field (Interaction, src/Interaction.cpp:589-860)
# ---------------------------------------------------------------------------
def read_points_file(path):
    """Spherical integration points as Interaction::generateWeights reads them (src/Interaction.cpp:744-764):
    one point per line, x y z [w]; the direction is normalised, the weight is the fourth column or 4 pi when
    there is none, divided by the number of points.  Returns an (n, 4) array (ux, uy, uz, weight).  Blank lines
    are skipped."""
    rows = []
    with open(path) as f:
        for line in f:
            vals = [float(v) for v in line.split()]
            if not vals:
                continue
            if len(vals) < 3:
                raise ValueError(f"{path}: a point needs three coordinates: {line!r}")
            vals.append(4 * np.pi)
            r = np.sqrt(vals[0] ** 2 + vals[1] ** 2 + vals[2] ** 2)
            rows.append([vals[0] / r, vals[1] / r, vals[2] / r, vals[3]])
    pts = np.array(rows, dtype=np.float64).reshape(-1, 4)
    pts[:, 3] /= len(rows)
    return pts


def _cxx(x):
    """A double as a default-formatted C++ stream prints it (6 significant digits, %g)."""
    return "%g" % x


def potential_identifier(kind, sigma, eps, rcut, alpha=50.0, r_att_min=0.0, bh=False):
    """Potential1::getIdentifier (include/Potential1.h:160-165, 227-232, 297-302): note that rcut and the WCA
    limit are printed without a separator, as the reference does."""
    tail = _cxx(rcut) + _cxx(r_att_min) + "_" + str(int(bool(bh)))
    if kind == 1:
        return "TWF_" + _cxx(sigma) + "_" + _cxx(eps) + "_" + _cxx(alpha) + "_" + tail
    if kind == 2:
        return "WHDF_" + _cxx(sigma) + "_" + _cxx(eps) + "_" + tail
    return "LJ_" + _cxx(sigma) + "_" + _cxx(eps) + "_" + tail


def weight_cache_name(seq1, seq2, shape, identifier):
    """Interaction::readWeights' file name (src/Interaction.cpp:596-603)."""
    return f"weights_{seq1}_{seq2}_{shape[0]}_{shape[1]}_{shape[2]}_{identifier}_.dat"


def write_weight_cache(path, w, dx, identifier, points_file, precision=6):
    """The cache Interaction writes (src/Interaction.cpp:641-651, 856-858): "Nx Ny Nz dx" (dx in scientific
    notation with 16 digits after the point), the potential identifier, the points-file name, then the kernel before
    the division by kT in the DFT_Vec text format (6 significant digits, as the reference's stream prints)."""
    w = np.asarray(w, dtype=np.float64)
    with open(path, "w", newline="\n") as f:
        f.write(f"{w.shape[0]} {w.shape[1]} {w.shape[2]} {dx:.16e}\n")
        f.write(f"{identifier}\n")
        f.write(f"{points_file}\n")
    flat = w.ravel()
    fmt = "%." + str(int(precision)) + "g"
    with open(path, "a", newline="\n") as f:
        f.write(f"{flat.size}\n")
        for b in range(0, flat.size, 1 << 16):
            f.write("".join(fmt % x + " " for x in flat[b:b + (1 << 16)].tolist()))
        f.write("\n")


def read_weight_cache(path, shape, dx, identifier, points_file):
    """Interaction::readWeights (src/Interaction.cpp:589-656): the kernel (before / kT) if the file exists and its
    grid, spacing (relative 1e-8), identifier and points-file name match, else None."""
    try:
        f = open(path)
    except OSError:
        return None
    with f:
        head = f.readline().split()
        ident = f.readline().strip()
        pfile = f.readline().rstrip("\n")
        if len(head) < 4:
            return None
        nx, ny, nz, d = int(head[0]), int(head[1]), int(head[2]), float(head[3])
        if (nx, ny, nz) != tuple(shape) or abs(dx - d) > 1e-8 * (dx + d):
            return None
        tok = ident.split()[0] if ident.split() else ""
        if tok != identifier:
            return None
        if pfile != points_file:
            return None
        n = int(f.readline().split()[0])
        vals = np.array(f.read().split(), dtype=np.float64)
    if vals.size != n or n != nx * ny * nz:
        return None
    return vals.reshape(shape)
