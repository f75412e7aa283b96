"""Closed forms that pin the oracle's boundary hooks and absorbing layer.

Each is written from the mathematics the reference's code implements, in a
form different from the code's own (vector identities instead of its
component-by-component expressions), so that a transposed operand, a swapped
width or shift, or a wrong Yee offset in the oracle shows up as a mismatch.
"""
import math

import numpy as np

# Yee positions of the six components, offsets from the cell corner (i, j, k)
# (user_psr.F90:157-159, :219-221, :246-248, :273-275; field_solver.F90:225-260)
YEE = {
    "ex": (0.5, 0.0, 0.0), "ey": (0.0, 0.5, 0.0), "ez": (0.0, 0.0, 0.5),
    "bx": (0.0, 0.5, 0.5), "by": (0.5, 0.0, 0.5), "bz": (0.5, 0.5, 0.0),
}


def dipole(r, radius, angle, phase):
    """Point dipole of moment radius^3 (sin a cos p, sin a sin p, cos a) at
    offset r (..., 3): B = (3 (m.n) n - m) / |r|^3."""
    m = radius ** 3 * np.array([math.sin(angle) * math.cos(phase), math.sin(angle) * math.sin(phase),
                                math.cos(angle)])
    d = np.linalg.norm(r, axis=-1, keepdims=True)
    n = r / d
    return (3.0 * n * (n @ m)[..., None] - m) / d ** 3


def tanh_step(rad, rad0):
    """The conducting sphere's profile: 1 deep inside, 0 far outside, width 1
    in its argument (user_psr.F90:376-383)."""
    return 0.5 * (1.0 - np.tanh(rad - rad0))


def psr_blend(c_e, c_b, pos, radius, angle, omega, step, cc, e_widths=(0.25, 0.5), b_widths=(0.5, 0.25),
              e_shift=0.0, b_shift=4.0):
    """What the rotator's hook does to uniform fields E = c_e, B = c_b at
    offsets `pos` (..., 3) from the star's centre: every interpolated
    neighbour average then equals the uniform value, and the hook reduces to

      B' = B_d + (1 - s_r^B) [(C - B_d).r^] r^ + (1 - s_t^B) [(C - B_d) - ((C - B_d).r^) r^]
      E' = E_d + (1 - s_t^E) [(C - E_d) - ((C - E_d).r^) r^] + (1 - s_r^E) [(C - E_d).r^] r^

    B_d the dipole rotated to phase omega * step, E_d = -(Omega z^ x r) x B_d / c
    the corotation field; the radial (r) and tangential (t) parts relax with
    their own widths: B radial 1/2, B tangential 1/4 about R - 4; E
    tangential 1/2, E radial 1/4 about R (user_psr.F90:139-142).

    (e_widths, b_widths) = (radial, tangential) widths and the shifts below
    R, overridable to show that a test tells the right values from wrong
    ones.  Returns (E', B') at every offset.
    """
    d = np.linalg.norm(pos, axis=-1, keepdims=True)
    rh = pos / d
    bd = dipole(pos, radius, angle, omega * step)
    v = omega * np.stack([-pos[..., 1], pos[..., 0], np.zeros_like(pos[..., 0])], axis=-1)
    ed = -np.cross(v, bd) / cc

    def relax(c, f, w_rad, w_tan, centre):
        dv = c - f
        rad = np.sum(dv * rh, axis=-1, keepdims=True) * rh
        s_r = tanh_step(d / w_rad, centre / w_rad)
        s_t = tanh_step(d / w_tan, centre / w_tan)
        return f + (1.0 - s_r) * rad + (1.0 - s_t) * (dv - rad)

    b_new = relax(np.asarray(c_b, float), bd, b_widths[0], b_widths[1], radius - b_shift)
    e_new = relax(np.asarray(c_e, float), ed, e_widths[0], e_widths[1], radius - e_shift)
    return e_new, b_new


def psr_expected(n, c_e, c_b, radius, angle, omega, step, cc):
    """The six interior arrays (k, j, i) after the rotator's hook on uniform
    fields, for a single meshblock of n^3 cells centred at n/2: cells whose
    integer corner lies inside r < 2R take the blend, every other cell keeps
    the uniform value."""
    xc = 0.5 * n
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    inside = (i - xc) ** 2 + (j - xc) ** 2 + (k - xc) ** 2 < (2.0 * radius) ** 2
    out = {}
    for name, (ox, oy, oz) in YEE.items():
        pos = np.stack([i + ox - xc, j + oy - xc, k + oz - xc], axis=-1)[inside]
        e_new, b_new = psr_blend(c_e, c_b, pos, radius, angle, omega, step, cc)
        comp = "xyz".index(name[1])
        a = np.full((n, n, n), (c_e if name[0] == "e" else c_b)[comp], dtype=float)
        a[inside] = (e_new if name[0] == "e" else b_new)[:, comp]
        out[name] = a
    return out, inside


def absorber_factor(n, buff, name, fixed):
    """The absorbing rk3Update's damping factor of component `name` at every
    interior cell (k, j, i) of an n^3 box (field_solver.F90:198-269): cells
    whose centre lies beyond r_abs - 1 multiply by 1 - delta^3, delta the
    component's distance past r_abs = n/2 - buff in units of buff.  The
    fixed mode measures y with j and caps delta at 1 (the layer's edge is
    then zero, and the cube's corners beyond n/2 stay zero instead of
    growing); the reference as coded measures y with i and leaves delta
    uncapped."""
    c = 0.5 * n
    rabs = c - buff
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    jg = j if fixed else i
    ox, oy, oz = YEE[name]
    r_cell = np.sqrt((i - c) ** 2 + (jg - c) ** 2 + (k - c) ** 2)
    r = np.sqrt((i + ox - c) ** 2 + (jg + oy - c) ** 2 + (k + oz - c) ** 2)
    delta = (r - rabs) / buff
    if fixed:
        delta = np.minimum(delta, 1.0)
    f = np.where(delta > 0.0, 1.0 - delta ** 3, 1.0)
    return np.where(r_cell > rabs - 1.0, f, 1.0)
