"""The seeded Alfvenic spectrum (problems.alfven_modes / spectrum_fields),
the host form of tff_init_spectrum.  Runs without a GPU.

* the inverse-FFT mode sum equals the mode sum written out term by term
  (cos/sin of the phase), cell by cell, on a small box with unequal sides;
* the forward-difference divergence of B vanishes to rounding, and the
  backward one does not (a curl built with a transposed offset would
  reverse this);
* rms|dB| = rms|E| = amp (Parseval over the half-space modes);
* the mode list is the half space of the ball 1 <= |k| <= kmax (no k
  together with -k, no k = 0), and 2 kmax >= n is rejected.
"""
import math

import numpy as np
import pytest

from helpers import interior

import problems


def _direct(n, modes, q, bg):
    nx, ny, nz = n
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    v = np.full((nz, ny, nx), float(bg))
    for m in modes:
        th = 2 * np.pi * (m[0] * i / nx + m[1] * j / ny + m[2] * k / nz)
        v += m[3 + 2 * q] * np.cos(th) - m[4 + 2 * q] * np.sin(th)
    return v


def test_fft_mode_sum_equals_direct_sum():
    n = (14, 12, 10)
    p = problems.alfven(n, seed=7, kmax=4, amp=0.3)
    modes, bg = p.init[1], p.init[2]
    for q, f in enumerate(("ex", "ey", "ez", "bx", "by", "bz")):
        want = _direct(n, modes, q, bg[q])
        assert np.max(np.abs(interior(p.fields[f]) - want)) < 1e-14, f
        ghost = p.fields[f].copy()
        ghost[2:-2, 2:-2, 2:-2] = 0.0
        assert not ghost.any(), f
    assert (p.fields["b0z"] == 1.0).all() and not p.fields["b0x"].any() and not p.fields["b0y"].any()


def test_divergence_free_and_amplitude():
    amp = 0.05
    p = problems.alfven(24, seed=5, kmax=4, amp=amp)
    bx, by = interior(p.fields["bx"]), interior(p.fields["by"])
    bz = interior(p.fields["bz"]) - 1.0
    fwd = (np.roll(bx, -1, 2) - bx) + (np.roll(by, -1, 1) - by) + (np.roll(bz, -1, 0) - bz)
    bwd = (bx - np.roll(bx, 1, 2)) + (by - np.roll(by, 1, 1)) + (bz - np.roll(bz, 1, 0))
    assert np.abs(fwd).max() < 1e-15
    assert np.abs(bwd).max() > 1e-3
    e = [interior(p.fields[f]) for f in ("ex", "ey", "ez")]
    assert not e[2].any()
    assert math.isclose(math.sqrt(np.mean(bx ** 2 + by ** 2 + bz ** 2)), amp, rel_tol=1e-13)
    assert math.isclose(math.sqrt(np.mean(e[0] ** 2 + e[1] ** 2)), amp, rel_tol=1e-13)


def test_mode_list_is_a_half_ball():
    kmax = 5
    m = problems.alfven_modes((16, 16, 16), seed=1, kmax=kmax)
    k = m[:, :3].astype(int)
    assert np.array_equal(k, m[:, :3])
    k2 = (k ** 2).sum(1)
    assert (k2 >= 1).all() and (k2 <= kmax * kmax).all()
    keys = {tuple(v) for v in k}
    assert len(keys) == len(k)
    assert not any(tuple(-v) in keys for v in k)
    # every lattice point of the ball but the origin is k or -k of one mode
    ball = sum(1 for a in range(-kmax, kmax + 1) for b in range(-kmax, kmax + 1) for c in range(-kmax, kmax + 1)
               if 1 <= a * a + b * b + c * c <= kmax * kmax)
    assert 2 * len(k) == ball
    with pytest.raises(ValueError):
        problems.alfven_modes((10, 16, 16), kmax=5)


def test_host_false_builds_no_arrays():
    for p in (problems.alfven(16, kmax=3, host=False), problems.loop(16, host=False),
              problems.psr(16, radius=3.0, r_lc=6.0, absorb_buff=4, host=False),
              problems.uniform(16, host=False)):
        assert p.fields is None and p.init is not None, p.name


def test_park_miller_is_the_minimal_standard():
    """The generator of the reference's randomNum (auxiliary.F90:246-257):
    from seed 1, the Lehmer sequence x <- 16807 x mod (2^31 - 1) reaches
    1043618065 at the 10000th draw (Park and Miller, CACM 31 (1988) 1192,
    the published check value); each draw returns x / 2^31; the reference's
    start is 123457 + rank."""
    g = problems.ParkMiller()
    assert g.seed == 123457.0
    g.seed = 1.0
    for _ in range(10000):
        u = g.uniform()
    assert g.seed == 1043618065.0
    assert u == 1043618065.0 / 2147483648.0
    # Box-Muller normals from it: mean ~ 0 and variance ~ 1 over many draws
    h = problems.ParkMiller(5)
    z = np.array([h.normal() for _ in range(20000)])
    assert abs(z.mean()) < 0.03 and abs(z.var() - 1.0) < 0.03
