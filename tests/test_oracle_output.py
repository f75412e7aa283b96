"""The oracle's output path (node interpolation and downsampling of
writeFields_hdf5) against the reference's own compiled interpolation, a
closed form and brute force."""
import numpy as np
import pytest

from oracle.oracle import NG, Oracle, RefSolver

E6 = ("ex", "ey", "ez", "bx", "by", "bz")


def _random_state(n, dtype, seed):
    rng = np.random.default_rng(seed)
    o = Oracle(n, dtype=dtype)
    for f in E6:
        o.field(f)[...] = rng.standard_normal(o.shape).astype(dtype)
    return o, rng


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_interp_bitwise_equals_reference(dtype):
    """orc_interp restates interpFromEdges / interpFromFaces
    (helpers.F90:93-257); the compiled reference routines give the same bits
    at random cells and offsets, including offset 0 (the output nodes)."""
    if not RefSolver.available(dtype):
        pytest.skip("oracle/_ref not built")
    n = (7, 6, 5)
    o, rng = _random_state(n, dtype, 11)
    r = RefSolver(n, dtype=dtype)
    for f in E6:
        r.put(f, o.field(f))
    offsets = [(0.0, 0.0, 0.0), (0.5, 0.25, 0.75)] + [tuple(rng.random(3)) for _ in range(6)]
    for dx, dy, dz in offsets:
        # the reference rounds the offsets to default REAL on entry
        dx, dy, dz = (float(np.dtype(dtype).type(v)) for v in (dx, dy, dz))
        for _ in range(12):
            i, j, k = (int(rng.integers(-1, s + 1)) for s in n)
            got = o.interp(dx, dy, dz, i, j, k).astype(np.float64)
            ref = r.interp(dx, dy, dz, i, j, k)
            assert np.array_equal(got, ref), ((dx, dy, dz), (i, j, k), got, ref)


def test_interp_of_linear_fields_is_exact():
    """Components of a linear field L sampled at their Yee positions (E on
    edges, B on faces) interpolate to L(i + dx, j + dy, k + dz): a wrong
    corner offset or term breaks it (dyadic values, so FP64 is exact)."""
    n = 6
    o = Oracle(n)
    k, j, i = np.meshgrid(*(np.arange(-NG, n + NG, dtype=np.float64),) * 3, indexing="ij")
    coef = {"ex": (0.5, 1.0, -2.0, 0.25), "ey": (-1.0, 0.75, 0.5, 2.0), "ez": (2.0, -0.25, 1.5, -1.0),
            "bx": (1.0, 0.5, 0.25, -0.5), "by": (-0.5, 2.0, -1.0, 0.75), "bz": (0.25, -1.5, 1.0, 0.5)}
    lin = lambda c, x, y, z: c[0] + c[1] * x + c[2] * y + c[3] * z
    # Yee positions: E_a at the cell's a-edge midpoint, B_a at its a-face centre
    pos = {"ex": (0.5, 0, 0), "ey": (0, 0.5, 0), "ez": (0, 0, 0.5),
           "bx": (0, 0.5, 0.5), "by": (0.5, 0, 0.5), "bz": (0.5, 0.5, 0)}
    for f, c in coef.items():
        px, py, pz = pos[f]
        o.field(f)[...] = lin(c, i + px, j + py, k + pz)
    for dx, dy, dz in [(0, 0, 0), (0.5, 0.25, 0.75), (0.125, 1.0, 0.5)]:
        for (ci, cj, ck) in [(0, 0, 0), (3, 2, 4), (-1, 5, 1), (6, -1, 6)]:
            got = o.interp(dx, dy, dz, ci, cj, ck)
            want = [lin(coef[f], ci + dx, cj + dy, ck + dz) for f in E6]
            assert np.array_equal(got, want), (f, (dx, dy, dz), (ci, cj, ck), got, want)


@pytest.mark.parametrize("istep", [2, 3, 4])
def test_output_sampling_against_brute_force(istep):
    """writeFields_hdf5 with output_istep >= 2 on one meshblock at the
    origin samples nodes 0, istep, ... below the meshblock size
    (write_output.F90:268-293): ceil(s / istep) of them per axis, each the
    offset-0 interpolation there."""
    n = (9, 8, 7)
    o, _ = _random_state(n, np.float32, 5)
    out = o.output_fields(istep)
    want_shape = tuple(-(-s // istep) for s in reversed(n))
    for f in E6:
        assert out[f].shape == want_shape
    for k1 in range(want_shape[0]):
        for j1 in range(want_shape[1]):
            for i1 in range(want_shape[2]):
                v = o.interp(0.0, 0.0, 0.0, i1 * istep, j1 * istep, k1 * istep)
                for q, f in enumerate(E6):
                    assert out[f][k1, j1, i1] == v[q]
