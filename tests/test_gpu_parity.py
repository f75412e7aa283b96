"""Parity of the HIP path (libtff.so through its C ABI) with the CPU oracle.

FP64 and FP32 follow the reference's operation order with fused multiply-add
disabled, so per-operation results are compared bit for bit; whole runs use
the FP64 gate of the measurement plan (normwise max-relative error <= 1e-12
after the run) where transcendental functions in the boundary hooks may
round differently on the GPU.  FP32-bulk is compared with the FP64 oracle on
the energy history (relative error <= 1e-5).
"""
import math

import numpy as np
import pytest

from helpers import FIELDS6, field_errors, interior, normwise_max_rel, random_fields
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

pytff = pytest.importorskip("pytff")
problems = pytest.importorskip("problems")


def _pair(n, boundary, cc=0.45, precision="fp64", absorb_buff=3, absorb_fixed=False):
    dt = np.float64 if precision == "fp64" else np.float32
    o = Oracle(n, boundary=boundary, cc=cc, absorb_buff=absorb_buff, absorb_fixed=absorb_fixed,
               dtype=np.float64 if precision == "fp32_bulk" else dt)
    g = pytff.Meshblock(n, boundary=boundary, cc=cc, precision=precision,
                        absorb_buff=absorb_buff, absorb_fixed=absorb_fixed)
    return o, g


def _compare_exact(o, g, names=FIELDS6, dtype=np.float64):
    for f in names:
        a = g.download(f, dtype=dtype)
        b = o.field(f)
        assert np.array_equal(a, b, equal_nan=True), (
            f"{f}: max diff {np.nanmax(np.abs(a.astype(np.float64) - b)):.3e}")


BOUNDARIES = [(1, 1, 1), (1, 1, 0), (0, 1, 0), (0, 0, 0), (2, 2, 2)]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("boundary", BOUNDARIES)
def test_reference_operations_bitwise(precision, boundary):
    """initRKstep, rk3Step (3 stage coefficient sets), exchangeFields and
    cleanEpar one by one, on random fields with every ghost filled: bit for
    bit equal to the oracle (itself bit for bit equal to field_solver.F90)."""
    n = (20, 14, 11)
    dt = np.float64 if precision == "fp64" else np.float32
    print(f"case precision={precision} boundary={boundary} n={n}", flush=True)
    o, g = _pair(n, boundary, precision=precision)
    rng = np.random.default_rng(7)
    fl = random_fields(o.shape, rng, dtype=dt)
    o.set_fields(**fl)
    g.set_fields(**fl)
    coef = pytff.rk_coefficients(g.precision)
    o.exchange_fields(True)
    g.exchangeFields(True)
    _compare_exact(o, g, FIELDS6 + ("b0x", "b0y", "b0z"), dtype=dt)
    o.init_rk_step()
    g.initRKstep()
    for c in coef:
        o.rk3_step(*c)
        g.rk3Step(*c)
        _compare_exact(o, g, dtype=dt)
        o.exchange_fields()
        g.exchangeFields()
        o.clean_epar()
        g.cleanEpar()
        _compare_exact(o, g, dtype=dt)
        o.exchange_fields()
        g.exchangeFields()
    _compare_exact(o, g, dtype=dt)


def _run_both(p, nsteps, precision="fp64", graph=False):
    dt = np.float32 if precision == "fp32" else np.float64
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc, absorb_buff=p.absorb_buff,
               absorb_fixed=p.absorb_fixed, dtype=dt)
    g = pytff.Meshblock(p.n, boundary=p.boundary, cc=p.cc, precision=precision,
                        absorb_buff=p.absorb_buff, absorb_fixed=p.absorb_fixed)
    fl = {k: v.astype(dt) for k, v in p.fields.items()}
    o.set_fields(**fl)
    o.set_bc(p.bc_kind, p.bc_params)
    g.set_fields(**fl)
    g.set_bc(p.bc_kind, p.bc_params)
    o.preloop()
    g.preloop()
    o.run(0, nsteps)
    if graph:
        g.run(0, nsteps)
    else:
        for s in range(nsteps):
            g.step(s)
    return o, g


SMALL = {
    "emwave": lambda: problems.emwave(32, mode=(1, 1), steps=20),
    "alfven": lambda: problems.alfven(24, seed=5, kmax=4, steps=20),
    "loop": lambda: problems.loop(32, steps=12, twist_angle=1.0),
    "loop_shipped": lambda: problems.loop(32, steps=12, twist=False),
    "loop_driven": lambda: problems.loop(32, steps=12, twist_angle=1.0, driver=problems.LOOP_DRIVEN),
    "psr": lambda: problems.psr(40, radius=5.0, r_lc=12.0, absorb_buff=6, periods=0.2),
    "psr_faithful": lambda: problems.psr(40, radius=5.0, r_lc=12.0, absorb_buff=6, periods=0.2,
                                         absorb_fixed=False),
    "uniform": lambda: problems.uniform(24, L0=4.0, steps=12),
}


@pytest.mark.parametrize("name", list(SMALL))
def test_fp64_steps_match_oracle(name):
    p = SMALL[name]()
    nsteps = min(p.steps, 12)
    print(f"case {name} n={p.n} steps={nsteps}", flush=True)
    o, g = _run_both(p, nsteps)
    err = field_errors(g.download, o.field)
    per_comp = {f: normwise_max_rel(g.download(f), o.field(f)) for f in FIELDS6}
    print(f"{name}: normwise max-rel E {err['E']:.3e} B {err['B']:.3e}; per component {per_comp}")
    assert max(err.values()) <= 1e-12
    if p.bc_kind in (0, 1) and not (p.bc_kind == 1 and p.bc_params[8]):
        # no transcendental function on the path: bit for bit
        _compare_exact(o, g)


@pytest.mark.parametrize("name", ["emwave", "alfven", "loop_shipped"])
def test_fp32_steps_bitwise(name):
    """The FP32 path reproduces the reference's shipped precision exactly."""
    p = SMALL[name]()
    o, g = _run_both(p, 8, precision="fp32")
    _compare_exact(o, g, dtype=np.float32)


@pytest.mark.parametrize("name", ["alfven", "loop", "psr"])
def test_graph_replay_equals_eager(name):
    p = SMALL[name]()
    ga = pytff.Meshblock(p.n, boundary=p.boundary, cc=p.cc, absorb_buff=p.absorb_buff,
                         absorb_fixed=p.absorb_fixed)
    gb = pytff.Meshblock(p.n, boundary=p.boundary, cc=p.cc, absorb_buff=p.absorb_buff,
                         absorb_fixed=p.absorb_fixed)
    for g in (ga, gb):
        problems.load(g, p)
        g.preloop()
    for s in range(7):
        ga.step(s)
    gb.run(0, 3)
    gb.run(3, 4)
    for f in FIELDS6:
        assert np.array_equal(ga.download(f), gb.download(f), equal_nan=True), f


def test_plane_wave_closed_form_gpu():
    """C1 form at 64^3, 200 steps on the GPU against the closed-form SSP-RK3
    growth of the discrete Yee eigenmode."""
    amp, mode = 1e-2, (1, 1)
    p = problems.emwave(64, amp=amp, mode=mode, cc=0.5, steps=200)
    g = pytff.Meshblock(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(g, p)
    g.preloop()
    g.run(0, 200)
    dbz = interior(g.download("bz")) - 1.0
    ref = problems.emwave_closed_form(p, amp, mode, 200)
    err = np.max(np.abs(dbz - ref)) / amp
    print(f"plane wave closed-form error {err:.3e}")
    assert err < 1e-10


def test_fp32_bulk_energy_vs_fp64_oracle():
    p = problems.alfven(32, seed=2, kmax=6, steps=40)
    o, g = _run_both(p, 0, precision="fp32_bulk")
    e0 = sum(o.reduce()[:2])
    eo, eg = [], []
    for s in range(40):
        o.step(s)
        g.step(s)
        eo.append(sum(o.reduce()[:2]))
        eg.append(sum(g.reduce()[:2]))
    eo, eg = np.array(eo), np.array(eg)
    rel = np.max(np.abs(eg - eo) / eo)
    drel = np.max(np.abs((eg - e0) - (eo - e0))) / max(np.max(np.abs(eo - e0)), 1e-300)
    print(f"fp32-bulk energy rel err {rel:.3e}, delta-energy rel err {drel:.3e}")
    assert rel <= 1e-5
    for f in FIELDS6:
        assert normwise_max_rel(g.download(f), o.field(f)) < 1e-4


@pytest.mark.parametrize("precision", ["fp32_bulk", "fp32"])
def test_psr_low_precision_finite_and_close_to_fp64(precision):
    """The rotator with the fixed absorbing layer: its corners beyond r_max
    are zeroed (B = 0) while E arrives there at underflowing amplitudes, the
    case that an |E|^2 reciprocal overflowing to +inf turns into 0 * inf =
    NaN.  Every field must stay finite and the energy close to the
    oracle's (FP64 for FP32-bulk, FP32 for FP32)."""
    p = problems.psr(64, radius=5.0, r_lc=20.0, absorb_buff=8)
    nsteps = 60
    o, g = _run_both(p, 0, precision=precision)
    o.run(0, nsteps)
    g.run(0, nsteps)
    for f in FIELDS6:
        assert np.isfinite(g.download(f)).all(), f
    eo, eg = sum(o.reduce()[:2]), sum(g.reduce()[:2])
    print(f"psr {precision}: energy {eg:.9e} vs oracle {eo:.9e}")
    assert abs(eg - eo) <= 1e-5 * eo


def test_reduce_matches_oracle():
    p = problems.alfven(24, seed=9, kmax=4)
    o, g = _run_both(p, 3)
    ro, rg = o.reduce(), g.reduce()
    for a, b in zip(rg, ro):
        assert abs(a - b) <= 1e-12 * max(abs(b), 1e-300) + 1e-300


def test_errors_are_returned_not_raised_in_c():
    g = pytff.Meshblock(8)
    with pytest.raises(pytff.TffError):
        g.rk3Step(1.0, 0.0, 1.0)  # before initRKstep: TFF_ESTATE
    with pytest.raises(pytff.TffError):
        g.set_bc(7, ())
    with pytest.raises(pytff.TffError):
        pytff.Meshblock(1)
    with pytest.raises(pytff.TffError):
        g.set_tuning(0, 65)  # z-chunks above 64 do not fit a tile's plane mask
    with pytest.raises(pytff.TffError):
        g.set_tuning(6, 8)   # no such tile shape


def test_psr_parameters_changed_between_graph_runs():
    """tff_set_bc between two graph runs on the per-procedure path (whose
    state is always clean): the second run must use the new sphere (a graph
    captured with the old staging buffer and box would replay stale
    parameters)."""
    p = SMALL["psr"]()
    o, g = _pair(p.n, p.boundary, cc=p.cc, absorb_buff=p.absorb_buff, absorb_fixed=p.absorb_fixed)
    for t in (o, g):
        problems.load(t, p)
        t.preloop()
    g.set_path(False)
    o.run(0, 3)
    g.run(0, 3)
    params = list(p.bc_params)
    params[0] = 8.0  # radius 5 -> 8
    o.set_bc(p.bc_kind, params)
    g.set_bc(p.bc_kind, params)
    o.run(3, 3)
    g.run(3, 3)
    err = field_errors(g.download, o.field)
    print("psr radius changed between runs:", err)
    assert max(err.values()) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_open_box_settled_ghosts_bitwise(precision):
    """Every axis open and no hook: after one full step the fused path stops
    carrying and cleaning the ghosts (they no longer change).  Ten steps of
    captured graphs equal the oracle bit for bit."""
    p = problems.alfven(24, seed=11, kmax=4, steps=10)
    p.boundary = (0, 0, 0)
    o, g = _run_both(p, 10, precision=precision, graph=True)
    _compare_exact(o, g, dtype=np.float32 if precision == "fp32" else np.float64)
