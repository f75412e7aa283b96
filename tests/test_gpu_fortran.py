"""The Fortran driver (tristan-ff_amd/fortran: deck parser, main loop in the
order of src/tristanmainloop.F90, user files) calling libtff.so through
ISO_C_BINDING, compared with the oracle run on the same problem."""
import os
import subprocess

import numpy as np
import pytest

from helpers import FIELDS6, field_errors
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
problems = pytest.importorskip("problems")

BUILD = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     "tristan-ff_amd", "_build")

EMWAVE_DECK = """#
#  plane wave over B0 = z
#
<node_configuration>
  sizex         = 1
<time>
  last          = {last}            # last timestep
<grid>
  mx0           = {n}
  my0           = {n}
  mz0           = {n}
  boundary_x    = 1
  boundary_y    = 1
  boundary_z    = 1
<output>
  start         = 0
  interval      = {interval}
  dump          = 1
<algorithm>
  c             = 0.5           # speed of light (dx/dt)
  fused         = {fused}
<problem>
  amplitude     = 1e-2
  kx            = 1
  ky            = 1
"""

LOOP_DECK = """<time>
  last          = {last}
<grid>
  mx0           = {n}
  my0           = {n}
  mz0           = {n}
  boundary_x    = 1
  boundary_y    = 1
  boundary_z    = 0
<output>
  interval      = {interval}
  dump          = 1
<algorithm>
  fused         = {fused}
<problem>
  a_L0          = 3.0
  h_L0          = 2.0
  L0            = {L0}
  r_L0          = 0.5
  twist         = 1
  beta_twist    = {beta}
  t0            = {t0}
"""


def _run(exe, deck, tmp_path, tag):
    path = os.path.join(BUILD, exe)
    if not os.path.exists(path):
        pytest.skip(f"{exe} not built")
    d = tmp_path / tag
    d.mkdir()
    (d / "input").write_text(deck)
    print(f"case {exe} {tag}", flush=True)
    r = subprocess.run([path, "-i", str(d / "input"), "-o", str(d / "out")], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    return d / "out", r.stdout


def _read_dump(path, n):
    a = np.fromfile(path, dtype=np.float64)
    shape = (6, n + 4, n + 4, n + 4)
    assert a.size == np.prod(shape)
    a = a.reshape(shape)
    return {f: a[i] for i, f in enumerate(FIELDS6)}


@pytest.mark.parametrize("fused", [1, 0])
def test_fortran_emwave_matches_oracle(tmp_path, fused):
    n, last = 24, 10
    out, _ = _run("tristan_ff_emwave", EMWAVE_DECK.format(n=n, last=last, interval=5, fused=fused),
                  tmp_path, f"fused{fused}")
    got = _read_dump(out / f"flds.{last:05d}.bin", n)
    p = problems.emwave(n, amp=1e-2, mode=(1, 1), cc=0.5)
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(o, p)
    o.preloop()
    o.run(0, last + 1)  # the step loop is inclusive: 0..last
    err = field_errors(lambda f: got[f], o.field)
    print("fortran emwave", err)
    assert max(err.values()) <= 1e-12


def test_fortran_loop_fused_equals_reference_order(tmp_path):
    n, last = 32, 6
    p = problems.loop(n, L0=2.5, twist=True, twist_angle=1.0, steps=last + 1)
    beta, t0 = p.bc_params[6], p.bc_params[7]
    outs = []
    for fused in (1, 0):
        out, _ = _run("tristan_ff_loop",
                      LOOP_DECK.format(n=n, last=last, interval=3, fused=fused, L0=2.5,
                                       beta=repr(beta), t0=repr(t0)),
                      tmp_path, f"loop{fused}")
        outs.append(_read_dump(out / f"flds.{last:05d}.bin", n))
    for f in FIELDS6:
        assert np.array_equal(outs[0][f], outs[1][f]), f
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(o, p)
    o.preloop()
    o.run(0, last + 1)
    err = field_errors(lambda f: outs[0][f], o.field)
    print("fortran loop", err)
    assert max(err.values()) <= 1e-12


def test_fortran_output_datasets_match_oracle(tmp_path):
    """With <output> istep = 2 the driver writes the six field datasets of
    writeFields_hdf5 from tff_output_fields; they equal the oracle's
    interpolation of the same dumped state (FP64, rounded to FP32 on output)."""
    n, last = 24, 10
    deck = EMWAVE_DECK.format(n=n, last=last, interval=5, fused=1).replace(
        "dump          = 1", "dump          = 1\n  istep         = 2")
    out, _ = _run("tristan_ff_emwave", deck, tmp_path, "istep2")
    state = _read_dump(out / f"flds.{last:05d}.bin", n)
    o = Oracle(n, boundary=(1, 1, 1), cc=0.5)
    o.set_fields(**state)
    want = o.output_fields(2)
    m = n // 2
    got = np.fromfile(out / f"flds.tot.{last:05d}.bin", dtype=np.float32)
    assert got.size == 6 * m ** 3
    got = got.reshape(6, m, m, m)
    for q, f in enumerate(("ex", "ey", "ez", "bx", "by", "bz")):
        assert np.array_equal(got[q], want[f].astype(np.float32)), f
