"""Pins of the CPU oracle (oracle/tff_oracle.c) -- runs without a GPU.

1. Against the reference itself: oracle/_ref is the reference's own
   src/algorithms/field_solver.F90 compiled from its sources; the oracle's
   initRKstep / rk3Step / cleanEpar must agree with it bit for bit, in FP64
   and in FP32 (the shipped precision), for every boundary mode.
2. Against closed forms and invariants for what the reference's compiled
   solver does not cover (stage order and RK coefficients, ghost fill, the
   user boundary hooks), each chosen so that a dropped term, wrong sign,
   wrong index or swapped operand fails it.
"""
import math
import os

import numpy as np
import pytest

from helpers import FIELDS6, NG, interior, normwise_max_rel, random_fields
from oracle.oracle import FIELDS, Oracle, RefSolver, rk_coefficients

import problems


# ---------------------------------------------------------------- 1. reference
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("boundary", [(1, 1, 1), (1, 1, 0), (0, 1, 0), (2, 2, 2)])
@pytest.mark.parametrize("n", [(12, 10, 14), (9, 16, 7)])
def test_oracle_bitwise_equals_reference_solver(dtype, boundary, n):
    if not RefSolver.available(dtype):
        pytest.skip("oracle/_ref not built (reference sources absent at build time)")
    o = Oracle(n, boundary=boundary, cc=0.45, absorb_buff=3, dtype=dtype)
    r = RefSolver(n, boundary=boundary, cc=0.45, absorb_buff=3, dtype=dtype)
    rng = np.random.default_rng(hash((n, boundary)) % 2 ** 32)
    fl = random_fields(o.shape, rng, dtype=dtype)
    # a few cells with |E| > |B| so the limiter branch runs
    fl["ex"][5, 5, 5] = 40.0
    fl["ey"][6, 4, 3] = -25.0
    for f, a in fl.items():
        o.field(f)[...] = a
        r.put(f, a)
    o.init_rk_step()
    r.init_rk_step()
    for c in rk_coefficients(dtype)[::3] and [rk_coefficients(dtype)[i:i + 3] for i in (0, 3, 6)]:
        o.rk3_step(*c)
        r.rk3_step(*c)
        o.clean_epar()
        r.clean_epar()
    for f in FIELDS:
        assert np.array_equal(o.field(f), r.get(f), equal_nan=True), f


def test_oracle_bitwise_equals_reference_golden_vectors():
    """The same pin from committed vectors (tests/golden/make_golden.py wrote
    them with the reference's compiled solver): holds where oracle/_ref is
    not built."""
    path = os.path.join(os.path.dirname(__file__), "golden", "ref_field_solver.npz")
    z = np.load(path)
    tags = sorted({k.split("/")[0] for k in z.files if not k.startswith("in_")})
    assert len(tags) == 6
    for tag in tags:
        name, bnd = tag.split("_")
        dtype = np.float64 if name == "f64" else np.float32
        boundary = tuple(int(c) for c in bnd)
        o = Oracle((6, 5, 4), boundary=boundary, cc=0.45, absorb_buff=2, dtype=dtype)
        for f in FIELDS:
            key = f"in_{bnd}/{f}"
            if key in z.files:
                o.field(f)[...] = z[key].astype(dtype)
        o.init_rk_step()
        co = rk_coefficients(dtype)
        for st in range(3):
            o.rk3_step(*co[3 * st:3 * st + 3])
            o.clean_epar()
        for f in FIELDS6:
            assert np.array_equal(o.field(f), z[f"{tag}/out/{f}"], equal_nan=True), (tag, f)


# ---------------------------------------------------------- 2. closed forms
@pytest.mark.parametrize("axis", ["z", "x", "y"])
@pytest.mark.parametrize("mode", [(1, 1), (2, 1), (0, 3)])
def test_plane_wave_closed_form(axis, mode):
    """Vacuum Yee eigenmode over a uniform background: the current vanishes
    (div_h E = 0 for the eigenvector) and cleanEpar is a no-op (E.B = 0 at
    every interpolation point), so every step multiplies the mode by the
    SSP-RK3 stability polynomial G(-i theta).  Pins the stage order, the three
    stage coefficient sets, the curl (all components over the three
    orientations) and the periodic ghost fill."""
    n, amp, steps = 16, 1e-3, 30
    p = problems.emwave(n, amp=amp, mode=mode, cc=0.45, axis=axis)
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(o, p)
    o.preloop()
    o.run(0, steps)
    comp = {"z": "bz", "x": "bx", "y": "by"}[axis]
    got = interior(o.field(comp)) - 1.0
    ref = problems.emwave_closed_form(p, amp, mode, steps, axis=axis)
    assert np.max(np.abs(got - ref)) / amp < 1e-11


def test_plane_wave_closed_form_detects_wrong_coefficient():
    """The same check fails when the last stage uses 1/2 instead of 2/3."""
    n, amp, steps = 16, 1e-3, 30
    p = problems.emwave(n, amp=amp, mode=(1, 1), cc=0.45)
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(o, p)
    o.preloop()
    o.coef = [1.0, 0.0, 1.0, 0.75, 0.25, 0.25, 1.0 / 3.0, 2.0 / 3.0, 0.5]
    o.run(0, steps)
    got = interior(o.field("bz")) - 1.0
    ref = problems.emwave_closed_form(p, amp, (1, 1), steps)
    assert np.max(np.abs(got - ref)) / amp > 1e-6


def test_div_b_conserved():
    """div_h curl_h = 0 and the RK combination is linear: div_h B stays at
    roundoff on a periodic box while the currents are strongly active."""
    p = problems.alfven(16, seed=4, amp=0.2, kmax=4)
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(o, p)
    o.preloop()
    d0 = o.reduce()[2]
    e0 = o.field("ex").copy()
    o.run(0, 15)
    assert d0 < 1e-14 and o.reduce()[2] < 1e-13
    assert np.max(np.abs(o.field("ex") - e0)) > 1e-3  # the fields did evolve


def test_uniform_field_is_static_bitwise():
    n = (10, 8, 12)
    o = Oracle(n, boundary=(1, 1, 1), cc=0.45)
    o.field("bx")[...] = 0.3
    o.field("by")[...] = -0.2
    o.field("bz")[...] = 1.0
    o.field("b0z")[...] = 1.0
    o.preloop()
    before = {f: o.field(f).copy() for f in FIELDS6}
    o.run(0, 5)
    for f in FIELDS6:
        assert np.array_equal(o.field(f), before[f]), f


def _linear_ex(alpha, bz=1.0, n=(24, 6, 6)):
    """E = (alpha x, 0, 0) at the Ex points x = i + 1/2, uniform B = bz z^, open
    boundaries; one rk3Step(0, 0, 1) leaves dU = L(U) in E, B."""
    o = Oracle(n, boundary=(0, 0, 0), cc=0.45)
    i = np.arange(-NG, n[0] + NG)[None, None, :]
    o.field("ex")[...] = alpha * (i + 0.5)
    o.field("bz")[...] = bz
    o.init_rk_step()
    o.rk3_step(0.0, 0.0, 1.0)
    return o, n


def test_drift_current_closed_form():
    """rho = alpha at every node; the only current is jy = -CC rho Ex B/(B^2+TINY)
    with Ex interpolated to the Ey point (= alpha i), so dEy = CC alpha^2 i B/(B^2+TINY)
    and every other component of dU is zero."""
    alpha, cc, tiny = 0.01, 0.45, 1e-7
    o, n = _linear_ex(alpha)
    i = np.arange(n[0])[None, None, :] + 0 * np.zeros((n[2], n[1], 1))
    ey = interior(o.field("ey"))
    expect = cc * alpha * alpha * i * 1.0 / (1.0 + tiny)
    assert np.allclose(ey, expect, rtol=1e-13, atol=1e-18)
    for f in ("ex", "ez", "bx", "by", "bz"):
        assert np.max(np.abs(interior(o.field(f)))) < 1e-15, f


def test_drift_current_limiter_closed_form():
    """With |E| > |B| the limiter rescales the interpolated transverse E to
    |B|: dEy = CC alpha sign(i) B^2/(B^2+TINY) where alpha |i| > B."""
    alpha, cc, tiny, b = 0.5, 0.45, 1e-7, 1.0
    o, n = _linear_ex(alpha, bz=b)
    ey = interior(o.field("ey"))[2, 2, :]
    i = np.arange(n[0])
    expect = np.where(alpha * i > b, cc * alpha * np.sign(i) * b * b / (b * b + tiny),
                      cc * alpha * alpha * i * b / (b * b + tiny))
    assert np.allclose(ey, expect, rtol=1e-13, atol=1e-18)


def test_clean_removes_parallel_e_uniform():
    """Uniform E and B: the interpolants are exact, so cleanEpar gives
    E - (E.B) B / |B|^2 -- E.B = 0 afterwards and E_perp unchanged."""
    n = (6, 6, 6)
    o = Oracle(n, boundary=(1, 1, 1))
    e = np.array([0.3, -0.7, 0.2])
    b = np.array([0.5, 0.4, -1.1])
    for c, name in enumerate("xyz"):
        o.field("e" + name)[...] = e[c]
        o.field("b" + name)[...] = b[c]
    o.clean_epar()
    got = np.array([interior(o.field("e" + c))[0, 0, 0] for c in "xyz"])
    expect = e - e.dot(b) * b / b.dot(b)
    assert np.allclose(got, expect, rtol=1e-14, atol=1e-16)
    assert abs(got.dot(b)) < 1e-15


def test_clean_nan_at_exact_null():
    """cleanEpar divides by |B|^2 with no floor (field_solver.F90:300): an exact
    null gives NaN (0/0), as in the reference."""
    o = Oracle((6, 6, 6), boundary=(1, 1, 1))
    o.field("ex")[...] = 0.0
    o.clean_epar()
    assert np.isnan(interior(o.field("ex"))).all()


@pytest.mark.parametrize("boundary", [(1, 1, 1), (1, 1, 0), (0, 1, 0), (2, 2, 2)])
def test_exchange_is_periodic_wrap(boundary):
    """One meshblock: a ghost cell takes the wrapped interior value iff every
    axis it lies outside of is periodic; all other ghosts are untouched."""
    n = (7, 9, 6)
    o = Oracle(n, boundary=boundary)
    rng = np.random.default_rng(0)
    a = rng.standard_normal(o.shape)
    o.field("ey")[...] = a
    o.exchange_fields()
    got = o.field("ey")
    wrapped = np.pad(interior(a), NG, mode="wrap")
    kk, jj, ii = np.meshgrid(*(np.arange(s) - NG for s in o.shape), indexing="ij")
    outside = [(ii < 0) | (ii >= n[0]), (jj < 0) | (jj >= n[1]), (kk < 0) | (kk >= n[2])]
    filled = np.ones(o.shape, bool)
    for ax in range(3):
        if boundary[ax] != 1:
            filled &= ~outside[ax]
    assert np.array_equal(got[filled], wrapped[filled])
    assert np.array_equal(got[~filled], a[~filled])


def test_loop_hook_floor_copy():
    p = problems.loop(16, twist=False)
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(o, p)
    rng = np.random.default_rng(1)
    for f in ("ex", "ey", "ez", "bz"):
        o.field(f)[...] = rng.standard_normal(o.shape)
    ez0 = o.field("ez").copy()
    o.user_bc(3)
    for f in ("ex", "ey", "bz"):
        a = o.field(f)
        assert np.array_equal(a[0], a[NG]) and np.array_equal(a[1], a[NG])
        assert np.array_equal(a[-1], a[-NG - 1]) and np.array_equal(a[-2], a[-NG - 1])
    assert np.array_equal(o.field("ez"), ez0)


def test_loop_driven_hook_closed_form():
    """Driven loop: on z <= 0 (the floor plane included) Ex is the footpoint
    driver -(x+1/2-xc) e^{-r1/r0} beta CC/r0 Bz_loc tanh(t/t0) with r1 the
    3D distance to source 1 and Bz_loc the potential field there; Bz is the
    potential field on the floor ghosts and untouched on the floor plane;
    the top ghost planes are a conductor (Ex = Ey = 0, Bz = the potential
    field); interior planes above the floor and Ez are untouched."""
    n = 16
    p = problems.loop(n, twist_angle=1.0, steps=40, driver=problems.LOOP_DRIVEN)
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(o, p)
    rng = np.random.default_rng(2)
    for f in ("ex", "ey", "ez"):
        o.field(f)[...] = rng.standard_normal(o.shape)
    before = {f: o.field(f).copy() for f in ("ex", "ey", "ez", "bz")}
    step = 9
    o.user_bc(step)
    xc, yc1, yc2, zc, bnorm, r0, beta, t0, mode = p.bc_params
    assert mode == problems.LOOP_DRIVEN
    i = np.arange(-NG, n + NG)[None, :].astype(float)
    j = np.arange(-NG, n + NG)[:, None].astype(float)

    def bz_pot(rx, ry1, ry2, rz):
        r1 = np.sqrt(rx * rx + ry1 * ry1 + rz * rz)
        r2 = np.sqrt(rx * rx + ry2 * ry2 + rz * rz)
        return bnorm * (rz / r1 ** 3 - rz / r2 ** 3), r1

    for kk, z in ((0, -2.0), (1, -1.0), (NG, 0.0)):
        rx, ry1, ry2, rz = i + 0.5 - xc, j - yc1, j - yc2, z - zc
        bzl, r1 = bz_pot(rx, ry1, ry2, rz)
        ex = -rx * np.exp(-r1 / r0) * beta * p.cc / r0 * bzl * math.tanh(step / t0)
        assert np.allclose(o.field("ex")[kk], ex, rtol=1e-12, atol=1e-300)
        bz, _ = bz_pot(i + 0.5 - xc, j + 0.5 - yc1, j + 0.5 - yc2, z - zc)
        if z < 0:
            assert np.allclose(o.field("bz")[kk], bz, rtol=1e-12)
        else:
            assert np.array_equal(o.field("bz")[kk], before["bz"][kk])
    for kk, z in ((-2, float(n)), (-1, float(n + 1))):
        assert not o.field("ex")[kk].any() and not o.field("ey")[kk].any()
        bz, _ = bz_pot(i + 0.5 - xc, j + 0.5 - yc1, j + 0.5 - yc2, z - zc)
        assert np.allclose(o.field("bz")[kk], bz, rtol=1e-12)
    for f in ("ex", "ey", "bz"):
        assert np.array_equal(o.field(f)[NG + 1:-NG], before[f][NG + 1:-NG])
    assert np.array_equal(o.field("ez"), before["ez"])


def test_uniform_hook_closed_form():
    """user_uniform: Ex = -(x+1/2-xc) e^{-r/r0} beta CC/r0 bnorm tanh(t/t0) on z <= 0."""
    p = problems.uniform(16, L0=4.0, beta_twist=0.3, t0=5.0)
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(o, p)
    step = 7
    o.user_bc(step)
    xc, yc, r0, beta, t0, bnorm = p.bc_params
    i = np.arange(-NG, 16 + NG)[None, :]
    j = np.arange(-NG, 16 + NG)[:, None]
    rx, ry = i + 0.5 - xc, j - yc
    rr = np.sqrt(rx * rx + ry * ry + 1e-7)
    ex = -rx * np.exp(-rr / r0) * beta * p.cc / r0 * bnorm * math.tanh(step / t0)
    for k in (0, 1, NG):  # z = -2, -1, 0
        assert np.allclose(o.field("ex")[k], ex, rtol=1e-13, atol=1e-16)
    assert np.all(o.field("ex")[NG + 1] == 0.0)


def test_psr_hook_limits():
    """Deep inside the star the hook sets B to the dipole and E to the
    corotation field -(Omega x r) x B / c; near r = 2R it leaves fields as
    they were (the tanh shapes saturate)."""
    p = problems.psr(64, radius=12.0, r_lc=40.0, absorb_buff=6)
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc, absorb_buff=6, absorb_fixed=True)
    problems.load(o, p)
    rng = np.random.default_rng(2)
    pert = {f: 1e-3 * rng.standard_normal(o.shape) for f in FIELDS6}
    for f in FIELDS6:
        o.field(f)[...] += pert[f]
    before = {f: o.field(f).copy() for f in FIELDS6}
    o.user_bc(0)
    R, ang, om, xc, yc, zc = p.bc_params
    # a cell at the centre + (1, 0, 0): r ~ 1 << R - 4
    i, j, k = 33, 32, 32
    c = (k + NG, j + NG, i + NG)
    bdx, _, _ = problems.dipole_field(p.bc_params, i, j + 0.5, k + 0.5)
    assert abs(o.field("bx")[c] - bdx) < 1e-8 * abs(bdx)
    bdx, bdy, bdz = problems.dipole_field(p.bc_params, i + 0.5, j, k)
    ry, rx = j - yc, i + 0.5 - xc
    ex_dip = -(om * rx * bdz) / p.cc
    assert abs(o.field("ex")[c] - ex_dip) < 1e-8 * abs(ex_dip) + 1e-12
    # a cell at radius 23 (< 2R = 24): the shapes vanish there
    i = 32 + 23
    c = (32 + NG, 32 + NG, i + NG)
    for f in FIELDS6:
        assert abs(o.field(f)[c] - before[f][c]) < 1e-6 * (abs(before[f][c]) + 1e-3)


def test_loop_driven_floor_conserves_div_b():
    """Every B component of the floor cells follows Faraday's law under the
    driver's E (bz of the floor plane is not pinned), so div_h B of every
    interior cell, the floor cells included, stays at its t = 0 value to
    roundoff while the footpoints turn.  Pinning bz at z = 0 (as
    user_uniform.F90:134-137 does for its own driver) changes it there by
    3.5 % of max|B| within these 40 steps."""
    n = 24
    p = problems.loop(n, twist_angle=3.0, steps=40, t0=2.0, driver=problems.LOOP_DRIVEN)
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(o, p)
    o.preloop()

    def div_h_b():
        bx, by, bz = (o.field(f) for f in ("bx", "by", "bz"))
        s = slice(NG, NG + n)
        s1 = slice(NG + 1, NG + n + 1)
        return (bx[s, s, s1] - bx[s, s, s]) + (by[s, s1, s] - by[s, s, s]) + (bz[s1, s, s] - bz[s, s, s])

    d0 = div_h_b()
    bz0 = o.field("bz")[NG].copy()
    o.run(0, 40)
    # the floor's bz moved (the driver acts), and div_h B did not
    assert np.abs(o.field("bz")[NG] - bz0).max() > 1e-3 * np.abs(bz0).max()
    scale = max(np.abs(o.field(f)).max() for f in ("bx", "by", "bz"))
    assert np.abs(div_h_b() - d0).max() <= 1e-12 * scale
