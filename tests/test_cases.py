"""Properties of the parity-test workloads that the GPU tests rely on."""
import sys
import os

import numpy as np

from helpers import near_null_cells

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tristan-ff_amd"))
import problems  # noqa: E402


def test_loop_cases_hold_near_null_cells():
    """The loop workloads of the fused-path bitwise tests start with cells at
    which FP32-bulk redoes the clean and currents in FP64.  The fused stage
    kernel defers those planes to its fix-up kernel, so the bitwise equality
    with the unfused path (which redoes inline) covers the deferral."""
    for p in (problems.loop(40, twist_angle=1.0, steps=12),
              problems.loop(36, twist=False, steps=12)):
        n = int(near_null_cells(p.fields).sum())
        assert n > 0, p.name


def test_batch_case_holds_near_null_cells():
    """So does the scan of the batched-path bitwise test (its loop
    instances), which covers the batched fix-up kernel."""
    probs = problems.scan(count=6, nmin=18, nmax=40, seed=3, steps=20)
    assert sum(int(near_null_cells(p.fields).sum()) for p in probs) > 0


def test_near_null_cells_of_a_uniform_field():
    """A uniform B has no near-null cell; zeroing B in a slab makes every
    cell of the slab one (|B|^2 below 1e-30)."""
    f = {c: np.zeros((12, 12, 12)) for c in ("bx", "by", "bz")}
    f["bz"][...] = 1.0
    assert not near_null_cells(f).any()
    f["bz"][5:7, :, :] = 0.0  # interior planes k = 3, 4
    near = near_null_cells(f)
    assert near[3:5].all()
    # the Ez point of plane 2 averages bz over k = 2, 3: still |B| = 1/2
    assert not near[:2].any() and not near[6:].any()
