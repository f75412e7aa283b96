"""The fused fast path (one kernel per stage, cleanEpar applied on the fly)
against the per-procedure path and the oracle.  The fused kernel evaluates
the same per-cell expressions, so the two paths agree bit for bit; results
are also independent of the tile shape and z-chunk the autotuner picks."""
import numpy as np
import pytest

from helpers import FIELDS6, field_errors, near_null_cells
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
pytff = pytest.importorskip("pytff")
problems = pytest.importorskip("problems")

CASES = {
    "alfven": lambda: problems.alfven(40, seed=5, kmax=6, amp=0.2),
    "emwave_odd": lambda: problems.emwave(36, mode=(2, 1)),
    "loop": lambda: problems.loop(40, twist_angle=1.0, steps=12),
    "loop_shipped": lambda: problems.loop(36, twist=False, steps=12),
    "loop_driven": lambda: problems.loop(36, twist_angle=1.0, steps=12, driver=problems.LOOP_DRIVEN),
    "psr": lambda: problems.psr(40, radius=5.0, r_lc=12.0, absorb_buff=6, periods=0.2),
    "psr_faithful": lambda: problems.psr(40, radius=5.0, r_lc=12.0, absorb_buff=6, periods=0.2,
                                         absorb_fixed=False),
    "uniform": lambda: problems.uniform(24, L0=4.0, steps=12),
    # x, y periodic and z open with a hook that reads no field: the fused
    # path wraps x and y and stores only the z-ghost planes; an uneven box
    # with partial tiles in x and y and fewer planes than a z-chunk
    "uniform_flat": lambda: problems.uniform((30, 20, 6), L0=4.0, steps=12),
}


def _mb(p, precision, fused):
    g = pytff.Meshblock(p.n, boundary=p.boundary, cc=p.cc, precision=precision,
                        absorb_buff=p.absorb_buff, absorb_fixed=p.absorb_fixed)
    problems.load(g, p)
    g.preloop()
    g.set_path(fused)
    return g


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk", "fp32"])
@pytest.mark.parametrize("name", list(CASES))
def test_fused_equals_unfused_bitwise(name, precision):
    p = CASES[name]()
    print(f"case {name} {precision} n={p.n}", flush=True)
    a = _mb(p, precision, fused=False)
    b = _mb(p, precision, fused=True)
    a.run(0, 9)
    b.run(0, 4)   # graph replay, then
    b.step(4)     # an eager step on the raw state,
    b.run(5, 4)   # and replay again
    dt = np.float64 if precision == "fp64" else np.float32
    for f in FIELDS6:
        x, y = a.download(f, dt), b.download(f, dt)
        assert np.array_equal(x, y, equal_nan=True), (f, float(np.nanmax(np.abs(x - y))))
    print("tuning", b.get_tuning())


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
@pytest.mark.parametrize("name", ["loop_driven", "psr", "uniform"])
def test_multi_step_graphs_keep_the_step(name, precision):
    """tff_run replays graphs of several steps whose hooks read the step
    from a device counter plus their offset in the graph: 19 steps (an eager
    step, two 8-step graphs and one-step graphs) of step-dependent boundary
    drivers equal the per-procedure path, step by step."""
    p = CASES[name]()
    print(f"case {name} {precision} n={p.n}", flush=True)
    a = _mb(p, precision, fused=False)
    b = _mb(p, precision, fused=True)
    for s in range(19):
        a.step(s)
    b.run(0, 19)
    dt = np.float64 if precision == "fp64" else np.float32
    for f in FIELDS6:
        x, y = a.download(f, dt), b.download(f, dt)
        assert np.array_equal(x, y, equal_nan=True), (f, float(np.nanmax(np.abs(x - y))))


def _nulls_everywhere():
    """Bz alternating in sign from plane to plane on top of a weak (1e-4)
    Alfvenic field: the Ez point's average over two planes cancels in every
    cell, so every plane of every tile takes the near-null FP64 redo."""
    p = problems.alfven(64, seed=4, kmax=4, amp=1e-4)
    bz = p.fields["bz"]
    sign = np.where(np.arange(bz.shape[0]) % 2 == 0, 1.0, -1.0)[:, None, None]
    bz[2:-2, 2:-2, 2:-2] *= sign[2:-2]
    return p


@pytest.mark.parametrize("kchunk", [4, 8, 64])
def test_fix_up_of_every_plane(kchunk):
    """With every plane marked, the stage kernel computes nothing and the
    fix-up kernel recomputes whole z-chunks (at 64^3 and kchunk 64, a full
    64-bit plane mask); the result still equals the unfused path bit for
    bit."""
    p = _nulls_everywhere()
    assert near_null_cells(p.fields).mean() > 0.9
    a = _mb(p, "fp32_bulk", fused=False)
    b = _mb(p, "fp32_bulk", fused=True)
    b.set_tuning(3, kchunk)
    a.run(0, 4)
    b.run(0, 4)
    for f in FIELDS6:
        x, y = a.download(f, np.float32), b.download(f, np.float32)
        assert np.isfinite(y).all(), f
        assert np.array_equal(x, y), (f, float(np.max(np.abs(x - y))))


@pytest.mark.parametrize("fused", [True, False])
def test_fp32_bulk_with_tiny_fields(fused):
    """Every field scaled by 1e-20 (|B|^2 ~ 1e-40, below the smallest normal
    FP32 number): FP32-bulk's cleanEpar rescales such B interpolants before
    its hardware reciprocal, so the run stays finite and follows FP64 (the
    projection does not depend on the scale of B)."""
    p = problems.alfven(24, seed=2, kmax=4, amp=0.2)
    for a in p.fields.values():
        a *= 1e-20
    ref = _mb(p, "fp64", fused)
    got = _mb(p, "fp32_bulk", fused)
    ref.run(0, 6)
    got.run(0, 6)
    for f in FIELDS6:
        assert np.isfinite(got.download(f)).all(), f
    err = field_errors(got.download, ref.download)
    print("tiny fields, fused" if fused else "tiny fields, unfused", err)
    assert max(err.values()) <= 1e-5


@pytest.mark.parametrize("name", ["alfven", "loop", "psr"])
def test_fused_fp64_matches_oracle(name):
    p = CASES[name]()
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc, absorb_buff=p.absorb_buff,
               absorb_fixed=p.absorb_fixed)
    problems.load(o, p)
    o.preloop()
    o.run(0, 10)
    g = _mb(p, "fp64", fused=True)
    g.run(0, 10)
    err = field_errors(g.download, o.field)
    print(name, err)
    assert max(err.values()) <= 1e-12


@pytest.mark.parametrize("name", ["alfven", "loop_driven", "psr"])
def test_fused_fp64_four_plane_chunks_match_oracle(name):
    """z-chunks of four planes (the autotuner's choice on small boxes): the
    prologue of every chunk is as long as the chunk itself, in each ghost
    mode (periodic, mixed, stored)."""
    p = CASES[name]()
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc, absorb_buff=p.absorb_buff,
               absorb_fixed=p.absorb_fixed)
    problems.load(o, p)
    o.preloop()
    o.run(0, 6)
    for tile in (0, 3):
        g = _mb(p, "fp64", fused=True)
        g.set_tuning(tile, 4)
        g.run(0, 6)
        err = field_errors(g.download, o.field)
        print(name, tile, err)
        assert max(err.values()) <= 1e-12


@pytest.mark.parametrize("name", ["alfven", "loop", "loop_driven", "uniform_flat"])
def test_results_independent_of_tile_shape(name):
    p = CASES[name]()
    ref = None
    for tile in range(6):
        for kc in (4, 8, 16, 64):
            g = _mb(p, "fp32_bulk", fused=True)
            try:
                g.set_tuning(tile, kc)
            except pytff.TffError:
                continue  # shape too large for this device's LDS per block
            g.run(0, 5)
            got = {f: g.download(f, np.float32) for f in FIELDS6}
            if ref is None:
                ref = got
            for f in FIELDS6:
                assert np.array_equal(got[f], ref[f], equal_nan=True), (tile, kc, f)


def test_reference_procedures_after_fused_steps():
    """Per-procedure calls on a state left raw by the fused path see the
    cleaned state (the library cleans it on demand)."""
    p = CASES["alfven"]()
    a = _mb(p, "fp64", fused=True)
    b = _mb(p, "fp64", fused=False)
    a.run(0, 3)
    b.run(0, 3)
    for g in (a, b):
        g.initRKstep()
        g.rk3Step(1.0, 0.0, 1.0)
        g.exchangeFields()
        g.cleanEpar()
        g.exchangeFields()
    for f in FIELDS6:
        assert np.array_equal(a.download(f), b.download(f), equal_nan=True), f
