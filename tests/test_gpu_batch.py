"""The batched path (one persistent stage kernel for many
uneven instances) against the same instances run one by one."""
import numpy as np
import pytest

from helpers import FIELDS6, field_errors
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
pytff = pytest.importorskip("pytff")
problems = pytest.importorskip("problems")


def _single(p, precision):
    g = pytff.Meshblock(p.n, boundary=p.boundary, cc=p.cc, precision=precision)
    problems.load(g, p)
    g.preloop()
    return g


@pytest.mark.parametrize("precision", ["fp32_bulk", "fp64"])
def test_batch_equals_single_instances_bitwise(precision):
    probs = problems.scan(count=6, nmin=18, nmax=40, seed=3, steps=20)
    print("case sizes", [p.n[0] for p in probs], [p.name for p in probs], flush=True)
    b = pytff.Batch(problems.batch_specs(probs), precision=precision)
    for q, p in enumerate(probs):
        problems.load(b[q], p)
        b[q].preloop()
    b.run(0, 4)
    b.run(4, 3)   # a second run continues from the state the first left
    print("batch tuning", b.get_tuning())
    dt = np.float64 if precision == "fp64" else np.float32
    for q, p in enumerate(probs):
        g = _single(p, precision)
        g.run(0, 7)
        for f in FIELDS6:
            x, y = b[q].download(f, dt), g.download(f, dt)
            assert np.array_equal(x, y, equal_nan=True), (p.name, f, float(np.nanmax(np.abs(x - y))))
    b.close()


def test_batch_fp64_matches_oracle():
    probs = problems.scan(count=3, nmin=20, nmax=28, seed=8, steps=10)
    b = pytff.Batch(problems.batch_specs(probs), precision="fp64")
    for q, p in enumerate(probs):
        problems.load(b[q], p)
        b[q].preloop()
    b.run(0, 8)
    for q, p in enumerate(probs):
        o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
        problems.load(o, p)
        o.preloop()
        o.run(0, 8)
        err = field_errors(b[q].download, o.field)
        print(p.name, err)
        assert max(err.values()) <= 1e-12
    b.close()


@pytest.mark.parametrize("precision", ["fp32_bulk", "fp64"])
def test_batch_of_every_ghost_mode_equals_single_instances(precision):
    """One batch holding an instance of each ghost mode of the stage kernel:
    the shipped loop (floor/top copy, stored ghosts), the ghost-only
    footpoint driver (stored), the driven loop (x/y wrapped, z-ghost planes
    stored) and a plane wave (all periodic); each equals its single-instance
    run bit for bit."""
    probs = [problems.loop(22, twist=False, steps=10),
             problems.loop(26, twist_angle=1.0, steps=10),
             problems.loop(30, twist_angle=1.0, steps=10, driver=problems.LOOP_DRIVEN),
             problems.emwave(20, mode=(1, 2), steps=10)]
    print("case", [(p.name, p.n[0], p.bc_params[8] if p.bc_params else None) for p in probs], flush=True)
    b = pytff.Batch(problems.batch_specs(probs), precision=precision)
    for q, p in enumerate(probs):
        problems.load(b[q], p)
        b[q].preloop()
    b.run(0, 6)
    dt = np.float64 if precision == "fp64" else np.float32
    for q, p in enumerate(probs):
        g = _single(p, precision)
        g.run(0, 6)
        for f in FIELDS6:
            x, y = b[q].download(f, dt), g.download(f, dt)
            assert np.array_equal(x, y, equal_nan=True), (p.name, f, float(np.nanmax(np.abs(x - y))))
    b.close()
