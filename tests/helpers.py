"""Shared helpers for the parity tests."""
import numpy as np

NG = 2
FIELDS6 = ("ex", "ey", "ez", "bx", "by", "bz")


def interior(a):
    return a[NG:-NG, NG:-NG, NG:-NG]


def normwise_max_rel(a, ref):
    """max|a - ref| / max|ref| (the normwise metric: pointwise relative error
    is undefined at zero crossings)."""
    d = np.max(np.abs(np.asarray(a, np.float64) - np.asarray(ref, np.float64)))
    s = np.max(np.abs(np.asarray(ref, np.float64)))
    return d / s if s > 0 else d


def random_fields(shape, rng, dtype=np.float64, b_offset=1.0):
    out = {}
    for f in FIELDS6 + ("b0x", "b0y", "b0z"):
        a = rng.standard_normal(shape)
        if f in ("bz", "b0z"):
            a = a + b_offset
        if f.startswith("e"):
            a = 0.3 * a
        out[f] = a.astype(dtype)
    return out


def field_errors(get, ref):
    """Normwise max-relative error of E and of B as vector fields:
    max_c max|U_c - U_ref,c| / max_c max|U_ref,c|.  (A component that is zero
    in exact arithmetic holds only roundoff, so normalising it by itself
    measures nothing but the rounding of transcendental functions.)"""
    out = {}
    for v in ("e", "b"):
        num = max(np.max(np.abs(np.asarray(get(v + c), np.float64) - np.asarray(ref(v + c), np.float64)))
                  for c in "xyz")
        den = max(np.max(np.abs(np.asarray(ref(v + c), np.float64))) for c in "xyz")
        out[v.upper()] = num / den if den > 0 else num
    return out


def near_null_cells(f):
    """Interior cells whose B interpolant at one of the three E points has
    |B|^2 below 1e-4 times the cell's own |B|^2: the cells at which the
    FP32-bulk kernels redo the clean and the currents in FP64.
    `f` maps bx, by, bz to arrays with ghosts, indexed [k][j][i]."""
    bx, by, bz = (np.asarray(interior(f[c]), np.float64) for c in ("bx", "by", "bz"))

    def at(a, di=0, dj=0, dk=0):  # a at (i + di, j + dj, k + dk), periodic
        return np.roll(a, (-dk, -dj, -di), axis=(0, 1, 2))

    bref = bx * bx + by * by + bz * bz
    # Ex point: bx over (i, i+1) x (j-1, j) x (k-1, k), by over k-1, k; bz over j-1, j
    x0 = 0.125 * sum(at(bx, di, dj, dk) for di in (0, 1) for dj in (-1, 0) for dk in (-1, 0))
    y0 = 0.5 * (at(by) + at(by, 0, 0, -1))
    z0 = 0.5 * (at(bz) + at(bz, 0, -1))
    # Ey point
    x1 = 0.5 * (at(bx) + at(bx, 0, 0, -1))
    y1 = 0.125 * sum(at(by, di, dj, dk) for di in (-1, 0) for dj in (0, 1) for dk in (-1, 0))
    z1 = 0.5 * (at(bz) + at(bz, -1))
    # Ez point
    x2 = 0.5 * (at(bx) + at(bx, 0, -1))
    y2 = 0.5 * (at(by) + at(by, -1))
    z2 = 0.125 * sum(at(bz, di, dj, dk) for di in (-1, 0) for dj in (-1, 0) for dk in (0, 1))
    near = np.zeros(bx.shape, bool)
    for m in (x0 * x0 + y0 * y0 + z0 * z0, x1 * x1 + y1 * y1 + z1 * z1, x2 * x2 + y2 * y2 + z2 * z2):
        near |= m < 1e-4 * bref
    return near


def h5_read(path, name, dtype, tmpdir):
    """One dataset of an HDF5 file, read with the HDF5 tools' h5dump (no
    Python HDF5 module here): raw little-endian values, flattened."""
    import os
    import shutil
    import subprocess
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    out = os.path.join(str(tmpdir), os.path.basename(str(path)) + "_" + name.replace(":", "_") + ".bin")
    subprocess.run([h5dump, "-d", name, "-b", "LE", "-o", out, str(path)], check=True, capture_output=True)
    return np.fromfile(out, dtype=dtype)


def h5_header(path):
    import shutil
    import subprocess
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    return subprocess.run([h5dump, "-H", str(path)], check=True, capture_output=True, text=True).stdout
