"""Closed-form pins of the oracle's boundary hooks and absorbing layer --
the parts of the oracle that the reference's compiled solver (oracle/_ref)
does not cover.  Runs without a GPU.

* the rotator's conducting-sphere blend (user_psr.F90:147-325) on uniform
  fields, against the vector form in closed_forms.psr_blend, over every cell
  of the supersphere including the transition shells where the radial and
  tangential widths differ; a blend with the widths or shifts swapped must
  fail the same comparison;
* the fixed absorbing layer (jg from j, delta capped at 1) and the layer as
  coded (jg from i, delta uncapped) on a uniform static field, whose update
  leaves exactly the damping factor;
* the uniform-twist footpoint driver (user_uniform.F90:114-142) through
  properties its definition fixes: rotation by 90 degrees about the box axis
  maps Ey onto Ex bit for bit (a radial field with the Yee offsets), the
  tanh(t/t0) ramp, the e^(-r/r0) profile, and linearity in beta.
"""
import math

import numpy as np
import pytest

from closed_forms import YEE, absorber_factor, psr_blend, psr_expected
from helpers import FIELDS6, NG, interior
from oracle.oracle import BC_PSR, Oracle, rk_coefficients

import problems

PSR = dict(n=40, radius=8.0, angle=math.radians(30.0), period=700.0, cc=0.5, step=37)
C_E = np.array([0.013, -0.021, 0.008])
C_B = np.array([0.7, -0.4, 1.3])


def _psr_oracle(n, radius, angle, period, cc, step):
    omega = 2.0 * math.pi / period
    o = Oracle(n, boundary=(2, 2, 2), cc=cc, absorb_buff=4, absorb_fixed=True)
    for f, c in zip(FIELDS6, np.concatenate([C_E, C_B])):
        o.field(f)[...] = c
    params = (radius, angle, omega, 0.5 * n, 0.5 * n, 0.5 * n)
    o.set_bc(BC_PSR, params)
    o.user_bc(step)
    return o, omega


def test_psr_blend_closed_form():
    """Rotator hook on uniform E, B: every component of every cell matches
    the radial/tangential relaxation toward the rotated dipole and the
    corotation field, with B's shells about R - 4 (radial width 1/2,
    tangential 1/4) and E's about R (tangential 1/2, radial 1/4)."""
    a = PSR
    o, omega = _psr_oracle(a["n"], a["radius"], a["angle"], a["period"], a["cc"], a["step"])
    exp, inside = psr_expected(a["n"], C_E, C_B, a["radius"], a["angle"], omega, a["step"], a["cc"])
    assert inside.sum() > 1000
    for f in FIELDS6:
        got = interior(o.field(f))
        ref = exp[f]
        scale = np.abs(ref) + np.abs(C_E if f[0] == "e" else C_B).max()
        err = np.max(np.abs(got - ref) / scale)
        print(f"{f}: max scaled error {err:.2e}")
        assert err < 1e-12, f
    # ghosts untouched
    for f, c in zip(FIELDS6, np.concatenate([C_E, C_B])):
        g = o.field(f).copy()
        g[NG:-NG, NG:-NG, NG:-NG] = c
        assert np.all(g == c), f


def test_psr_blend_closed_form_detects_swaps():
    """The same comparison fails for a blend with the radial and tangential
    widths exchanged (for E or for B) or with E's and B's shells moved onto
    each other: the cells of the transition shells tell them apart."""
    a = PSR
    o, omega = _psr_oracle(a["n"], a["radius"], a["angle"], a["period"], a["cc"], a["step"])
    n, xc = a["n"], 0.5 * a["n"]
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    inside = (i - xc) ** 2 + (j - xc) ** 2 + (k - xc) ** 2 < (2.0 * a["radius"]) ** 2
    args = (C_E, C_B)
    for name in ("ex", "bx"):
        ox, oy, oz = YEE[name]
        pos = np.stack([i + ox - xc, j + oy - xc, k + oz - xc], axis=-1)[inside]
        got = interior(o.field(name))[inside]
        col = 0 if name[0] == "e" else 1
        scale = np.max(np.abs(C_E if name[0] == "e" else C_B))
        blend = lambda **kw: psr_blend(*args, pos, a["radius"], a["angle"], omega, a["step"], a["cc"],
                                       **kw)[col][:, 0]
        assert np.max(np.abs(got - blend())) < 1e-12 * (np.max(np.abs(got)) + scale)
        wrongs = ([dict(e_widths=(0.5, 0.25)), dict(e_shift=4.0)] if name[0] == "e"
                  else [dict(b_widths=(0.25, 0.5)), dict(b_shift=0.0)])
        for kw in wrongs:
            diff = np.max(np.abs(got - blend(**kw)))
            print(name, kw, f"{diff:.3e}")
            assert diff > 1e-3 * scale, (name, kw)


@pytest.mark.parametrize("fixed", [True, False])
def test_absorber_closed_form(fixed):
    """A uniform B with E = 0 is a fixed point of rk3Push (no curl, no
    current), so one absorbing rk3Update leaves exactly the damping factor
    times B.  Checked on every cell, and by hand at three cells: inside the
    layer's inner edge (factor 1), on the x axis in the layer, and at the
    cube's corner beyond r_max = n/2 (fixed: delta capped, factor 0; as
    coded: 1 - delta^3 far below -1)."""
    n, buff = 32, 6
    o = Oracle(n, boundary=(2, 2, 2), cc=0.45, absorb_buff=buff, absorb_fixed=fixed)
    cb = {"bx": 0.3, "by": -0.7, "bz": 1.1}
    for f in FIELDS6 + ("b0x", "b0y", "b0z"):
        o.field(f)[...] = cb.get(f, 0.0)
    o.init_rk_step()
    o.rk3_step(*rk_coefficients(np.float64)[0:3])
    for f in ("bx", "by", "bz"):
        want = absorber_factor(n, buff, f, fixed) * cb[f]
        assert np.allclose(interior(o.field(f)), want, rtol=1e-14, atol=0), f
    for f in ("ex", "ey", "ez"):
        assert not interior(o.field(f)).any()
    # by hand (c = 16, r_abs = 10): bx of cell (29, 16, 16) sits at
    # (29, 16.5, 16.5), |r| = sqrt(169.5) = 13.01922, delta = 0.503204,
    # factor 1 - 0.127417 = 0.872583 (fixed: j = 16; as coded jg = i = 29
    # puts the cell at |r| = 18.4 and delta = 1.40: factor -1.77)
    got = o.field("bx")[16 + NG, 16 + NG, 29 + NG] / cb["bx"]
    if fixed:
        assert abs(got - 0.872583) < 1e-6
    else:
        r = math.sqrt(13 ** 2 + 13.5 ** 2 + 0.5 ** 2)
        assert abs(got - (1 - ((r - 10) / 6) ** 3)) < 1e-12 and got < -1.5
    # cell (25, 16, 16): its centre is at r = 9 = r_abs - 1, not beyond
    if fixed:
        assert o.field("by")[16 + NG, 16 + NG, 25 + NG] == cb["by"]
    # corner (0, 0, 0): bz at (0.5, 0.5, 0), |r| = 26.87, delta = 2.81
    corner = o.field("bz")[NG, NG, NG] / cb["bz"]
    if fixed:
        assert corner == 0.0
    else:
        r = math.sqrt(15.5 ** 2 + 15.5 ** 2 + 16 ** 2)
        assert abs(corner - (1 - ((r - 10) / 6) ** 3)) < 1e-12 and corner < -20


def test_clean_guard_at_exact_null():
    """The fixed mode's cleanEpar leaves E unchanged where |B|^2 = 0 (the
    reference divides by zero there: NaN)."""
    o = Oracle((8, 8, 8), boundary=(2, 2, 2), cc=0.45, absorb_buff=2, absorb_fixed=True)
    rng = np.random.default_rng(4)
    e = {f: rng.standard_normal(o.shape) for f in ("ex", "ey", "ez")}
    o.set_fields(**e)
    o.clean_epar()
    for f in ("ex", "ey", "ez"):
        assert np.array_equal(interior(o.field(f)), interior(e[f])), f


def _uniform(step, beta=0.3, n=16, L0=4.0, t0=5.0):
    p = problems.uniform(n, L0=L0, beta_twist=beta, t0=t0)
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(o, p)
    o.user_bc(step)
    return o, p


def test_uniform_driver_is_a_rotation_field():
    """The driver's E is radial about the box's vertical axis: turning the
    plane by 90 degrees maps the Ey points onto the Ex points, and the
    driver's values there agree bit for bit (a swapped offset, axis or sign
    breaks this)."""
    o, p = _uniform(9)
    n = p.n[0]
    xc = int(p.bc_params[0])
    ex, ey = o.field("ex"), o.field("ey")
    checked = 0
    for kk in range(NG + 1):  # z = -2, -1, 0
        for j in range(-NG, n + NG):
            for i in range(-NG, n + NG):
                # Ex at (i+1/2, j): offset (a, b) = (i + 1/2 - xc, j - xc);
                # Ey at offset (-b, a): i' = xc - b, j' + 1/2 = xc + a
                ip, jp = xc - (j - xc), xc + (i - xc)
                if -NG <= ip < n + NG and -NG <= jp < n + NG:
                    assert ey[kk, jp + NG, ip + NG] == ex[kk, j + NG, i + NG], (kk, j, i)
                    checked += 1
    assert checked > 500
    assert np.abs(ex[:NG + 1]).max() > 0


def test_uniform_driver_ramp_profile_and_beta():
    """E(t) / E(t') = tanh(t/t0) / tanh(t'/t0) at every driven point, E = 0
    at t = 0; along the x axis (j = yc) Ex / (x - xc) falls as
    e^(-r/r0) with r0 = r_L0 * L0; E is linear in beta."""
    t0 = 5.0
    o3, p = _uniform(3, t0=t0)
    o8, _ = _uniform(8, t0=t0)
    o0, _ = _uniform(0, t0=t0)
    assert not o0.field("ex")[:NG + 1].any() and not o0.field("ey")[:NG + 1].any()
    a3, a8 = o3.field("ex")[:NG + 1], o8.field("ex")[:NG + 1]
    nz = a8 != 0
    assert np.allclose(a3[nz] / a8[nz], math.tanh(3 / t0) / math.tanh(8 / t0), rtol=1e-13)
    ob, _ = _uniform(8, beta=0.6, t0=t0)
    assert np.allclose(ob.field("ex")[:NG + 1], 2.0 * a8, rtol=1e-13, atol=0)
    # radial profile on the x axis
    xc, yc, r0 = p.bc_params[0], p.bc_params[1], p.bc_params[2]
    assert r0 == 4.0  # r_L0 = 1 times L0 = 4
    row = o8.field("ex")[NG, int(yc) + NG, :]
    i = np.arange(-NG, p.n[0] + NG)
    a = i + 0.5 - xc
    r = np.sqrt(a * a + 1e-7)
    g = row / a
    ratio = g[1:] / g[:-1]
    assert np.allclose(ratio, np.exp(-(r[1:] - r[:-1]) / r0), rtol=1e-12)


def test_uniform_driver_field_is_the_initial_field():
    """The hook's Bz below the floor and in the top ghosts is the initial
    potential field, and the top ghosts' E is zero (a conductor)."""
    o, p = _uniform(4)
    bz0 = p.fields["bz"]
    for kk in list(range(NG + 1)) + [-1, -2]:
        assert np.allclose(o.field("bz")[kk], bz0[kk], rtol=1e-14), kk
    for kk in (-1, -2):
        assert not o.field("ex")[kk].any() and not o.field("ey")[kk].any()
