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
    """A uniform B has no near-null cell, and neither has a slab of zero or
    of tiny B (the test is relative to each cell's own |B|^2)."""
    f = {c: np.zeros((12, 12, 12)) for c in ("bx", "by", "bz")}
    f["bz"][...] = 1.0
    assert not near_null_cells(f).any()
    f["bz"][5:7, :, :] = 0.0  # interior planes k = 3, 4
    near = near_null_cells(f)
    # plane 4: own bz = 0 and the Ez point averages k = 4, 5 (|B| = 1/2);
    # the Ex point's bz is 0 -> |Bbar|^2 = 0 with bref = 0: exact zero, not near
    assert not near[4].any()
    # plane 3: own bz = 0, every interpolant of the Ex/Ey points is 0 and
    # the Ez point averages k = 3, 4 (zero): all-zero stencil, not near
    assert not near[3].any()
    # planes 2 and 5: own bz = 1, the Ez point of plane 2 averages k = 2, 3
    # (|B| = 1/2): not near
    assert not near[:3].any() and not near[5:].any()
    f["bz"][5:7, :, :] = 1e-20
    assert not near_null_cells(f).any()
    # bz = +1 on interior plane 4 and -1 on plane 5: plane 4's Ez point
    # averages the two to zero against its own |B|^2 = 1
    f["bz"][6, :, :] = 1.0
    f["bz"][7, :, :] = -1.0
    assert near_null_cells(f)[4].all()
