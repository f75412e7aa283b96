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
