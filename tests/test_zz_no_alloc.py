"""tff_output_fields_noalloc (the output entry point that allocates no
device memory) against tff_output_fields, through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
pytff = pytest.importorskip("pytff")
problems = pytest.importorskip("problems")

F6 = ("ex", "ey", "ez", "bx", "by", "bz")


def _mb(p, precision):
    g = pytff.Meshblock(p.n, boundary=p.boundary, cc=p.cc, precision=precision,
                        absorb_buff=p.absorb_buff, absorb_fixed=p.absorb_fixed)
    problems.load(g, p)
    g.preloop()
    return g


def _same(g, istep, dtype):
    a = g.output_fields(istep, dtype=dtype)
    b = g.output_fields(istep, dtype=dtype, noalloc=True)
    for f in F6:
        assert a[f].shape == b[f].shape
        assert np.array_equal(a[f], b[f], equal_nan=True), (istep, dtype, f)
    return a["ex"].size


def test_noalloc_output_equals_output_on_two_calls():
    """Two calls in a row, the second with a smaller output, each equal bit
    for bit to tff_output_fields on the same state (the second call must not
    see the first's staging)."""
    g1 = _mb(problems.alfven(48, seed=2, kmax=4), "fp32_bulk")
    g1.run(0, 3)  # raw fused state: the call cleans it first
    n1 = _same(g1, 2, np.float32)
    g2 = _mb(problems.loop(30, steps=10, twist_angle=1.0, driver=problems.LOOP_DRIVEN), "fp64")
    g2.run(0, 2)
    n2 = _same(g2, 3, np.float64)
    print(f"nodes per dataset: {n1}, then {n2}", flush=True)
    assert n2 < n1
    # and the first handle again, after the smaller call
    _same(g1, 4, np.float32)


def test_noalloc_output_rejects_outputs_beyond_its_scratch():
    g = pytff.Meshblock(224, precision="fp32")  # istep 2: 112^3 nodes x 6 x 4 B > 24 MiB
    with pytest.raises(pytff.TffError):
        g.output_fields(2, noalloc=True)
    g.output_fields(3, noalloc=True)  # 75^3 nodes fit
