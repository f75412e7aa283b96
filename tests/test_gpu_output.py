"""The output path on the GPU (writeFields_hdf5's node interpolation and
downsampling) against the oracle, on states advanced by the fused path."""
import numpy as np
import pytest

from helpers import FIELDS6
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
pytff = pytest.importorskip("pytff")
problems = pytest.importorskip("problems")


def _state(p, precision, steps):
    g = pytff.Meshblock(p.n, boundary=p.boundary, cc=p.cc, precision=precision,
                        absorb_buff=p.absorb_buff, absorb_fixed=p.absorb_fixed)
    problems.load(g, p)
    g.preloop()
    g.run(0, steps)
    return g


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32_bulk"])
@pytest.mark.parametrize("case", ["alfven", "loop"])
def test_output_fields_equal_oracle(case, precision):
    """The interpolants the GPU samples equal the oracle's (interpFromEdges
    / interpFromFaces restated, pinned to the reference's compiled routines)
    bit for bit, computed from the same downloaded state, for istep 2 and 3
    on a box whose sizes are not multiples of istep."""
    p = (problems.alfven(29, seed=3, kmax=5, amp=0.2) if case == "alfven"
         else problems.loop(27, twist_angle=1.0, steps=6, driver=problems.LOOP_DRIVEN))
    print(f"case {case} {precision} n={p.n}", flush=True)
    g = _state(p, precision, 5)
    dt = np.float64 if precision == "fp64" else np.float32
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc, dtype=dt)
    o.set_fields(**{f: g.download(f, dt) for f in FIELDS6})
    for istep in (2, 3):
        got = g.output_fields(istep, dtype=dt)
        want = o.output_fields(istep)
        for f in ("ex", "ey", "ez", "bx", "by", "bz"):
            assert got[f].shape == want[f].shape
            assert np.array_equal(got[f], want[f], equal_nan=True), (istep, f)


def test_output_fields_host_types_and_errors():
    """An FP64 handle widened to FP32 output equals its FP64 output rounded;
    output_istep 1 is rejected."""
    p = problems.alfven(20, seed=1, kmax=4, amp=0.1)
    g = _state(p, "fp64", 2)
    a = g.output_fields(2, dtype=np.float64)
    b = g.output_fields(2, dtype=np.float32)
    for f in a:
        assert np.array_equal(a[f].astype(np.float32), b[f])
    with pytest.raises(pytff.TffError):
        g.output_fields(1)
