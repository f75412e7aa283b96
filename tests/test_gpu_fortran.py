"""The Fortran driver (tristan-ff_amd/fortran: deck parser, main loop in the
order of src/tristanmainloop.F90, user files) calling libtff.so through
ISO_C_BINDING, compared with the oracle run on the same problem; its output
files (params, flds.tot + XDMF, domain: write_output.F90) read back with
h5dump; its parameter echo (printParams) and timing report (makeReport)."""
import dataclasses
import math
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import FIELDS6, field_errors, h5_header, h5_read
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
  sizey         = 1
  sizez         = 1
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

LOOP_DECK = """<node_configuration>
  sizex         = 1
  sizey         = 1
  sizez         = 1
<time>
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
    """With <output> istep = 2 the driver writes flds.tot.NNNNN (NNNNN the
    output index: steps 0, 5, 10 give 0, 1, 2) whose six field datasets
    equal the oracle's interpolation of the same dumped state (FP64, rounded
    to FP32 on output); var1..var3 are zero (user_emwave has no userOutput)
    and xx, yy, zz the node coordinates."""
    n, last = 24, 10
    deck = EMWAVE_DECK.format(n=n, last=last, interval=5, fused=1).replace(
        "dump          = 1", "dump          = 1\n  istep         = 2")
    out, _ = _run("tristan_ff_emwave", deck, tmp_path, "istep2")
    state = _read_dump(out / f"flds.{last:05d}.bin", n)
    o = Oracle(n, boundary=(1, 1, 1), cc=0.5)
    o.set_fields(**state)
    want = o.output_fields(2)
    m = n // 2
    path = out / "flds.tot.00002"
    hdr = h5_header(path)
    assert sorted(re.findall(r'DATASET "(\w+)"', hdr)) == sorted(
        ["ex", "ey", "ez", "bx", "by", "bz", "var1", "var2", "var3", "xx", "yy", "zz"])
    assert hdr.count(f"SIMPLE {{ ( {m}, {m}, {m} ) /") == 12 and hdr.count("H5T_IEEE_F32LE") == 12
    for f in ("ex", "ey", "ez", "bx", "by", "bz"):
        assert np.array_equal(h5_read(path, f, np.float32, tmp_path), want[f].astype(np.float32).ravel()), f
    for v in ("var1", "var2", "var3"):
        assert not h5_read(path, v, np.float32, tmp_path).any()
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    for name, c in zip(("xx", "yy", "zz"), (i, j, k)):
        assert np.array_equal(h5_read(path, name, np.float32, tmp_path), (2 * c).astype(np.float32).ravel())
    xdmf = (out / "flds.tot.00002.xdmf").read_text()
    assert f'Dimensions="{m:10d}{m:10d}{m:10d}"' in xdmf and "flds.tot.00002:/zz" in xdmf
    # params: the timestep and the reference's keys only, in their types
    pp = out / "params.00002"
    assert h5_read(pp, "timestep", np.int32, tmp_path)[0] == last
    assert h5_read(pp, "grd:mx0", np.int32, tmp_path)[0] == n
    assert h5_read(pp, "tim:last", np.int32, tmp_path)[0] == last
    assert h5_read(pp, "out:istep", np.int32, tmp_path)[0] == 2
    assert h5_read(pp, "alg:c", np.float32, tmp_path)[0] == np.float32(0.5)
    assert h5_read(pp, "prb:amplitude", np.float32, tmp_path)[0] == np.float32(1e-2)
    names = re.findall(r'DATASET "([^"]+)"', h5_header(pp))
    assert "alg:fused" not in names and "out:dump" not in names and "cpu:sizex" in names
    dom = out / "domain.00002"
    assert [int(h5_read(dom, v, np.int32, tmp_path)[0]) for v in ("x0", "y0", "z0", "sx", "sy", "sz")] == \
        [0, 0, 0, n, n, n]


PSR_DECK = """# inputs/input.psr with the grid scaled to 48^3 and the star to R = 6
<node_configuration>
  sizex         = 8
  sizey         = 8
  sizez         = 6
<time>
  last          = {last}
<grid>
  mx0           = 48
  my0           = 48
  mz0           = 48
  boundary_x    = 0
  boundary_y    = 0
  boundary_z    = 0
<output>
  start         = 0
  interval      = {interval}
  istep         = 2
  dump          = 1
<algorithm>
  c             = 0.5
<problem>
  psr_radius  = 6.0
  psr_angle   = 30.0
  psr_period  = 1600.0
"""


def test_fortran_psr_deck_matches_oracle(tmp_path):
    """The rotator deck (inputs/input.psr's keys and values, grid scaled):
    outflow boundaries, the conducting-sphere hook every stage; the driver's
    state against the oracle, and its parameter echo."""
    last = 8
    out, stdout = _run("tristan_ff_psr", PSR_DECK.format(last=last, interval=4), tmp_path, "psr")
    p = problems.psr(48, radius=6.0, angle_deg=30.0, cc=0.5)
    params = list(p.bc_params)
    params[2] = 2.0 * problems.PSR_PI / 1600.0  # psr_omega = 2 M_PI / psr_period (user_psr.F90:30-31)
    p = dataclasses.replace(p, boundary=(0, 0, 0), bc_params=tuple(params))
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(o, p)
    o.preloop()
    o.run(0, last + 1)
    got = _read_dump(out / f"flds.{last:05d}.bin", 48)
    err = field_errors(lambda f: got[f], o.field)
    print("fortran psr", err)
    assert max(err.values()) <= 1e-12
    # printParams (initialize.F90:74-107): one right-justified line per key
    assert "== Simulation parameters" in stdout
    assert f"{'prb':>30}:{'psr_radius':>20}:{'6.00':>10}" in stdout
    assert f"{'cpu':>30}:{'sizez':>20}:{6:10d}" in stdout
    assert f"{'grd':>30}:{'boundary_x':>20}:{0:10d}" in stdout
    # var1 = rho of the written state at the nodes (user_psr.F90:386-398)
    ex, ey, ez = got["ex"], got["ey"], got["ez"]
    sl = np.s_[2:-2:2, 2:-2:2, 2:-2:2]
    sm = np.s_[2:-2:2, 2:-2:2, 1:-3:2]
    rho = (ex[sl] - ex[sm]) + (ey[sl] - ey[2:-2:2, 1:-3:2, 2:-2:2]) + (ez[sl] - ez[1:-3:2, 2:-2:2, 2:-2:2])
    var1 = h5_read(out / "flds.tot.00002", "var1", np.float32, tmp_path)
    assert np.array_equal(var1, rho.astype(np.float32).ravel())


UNIFORM_DECK = """<node_configuration>
  sizex         = 1
  sizey         = 1
  sizez         = 1
<time>
  last          = {last}
<grid>
  mx0           = 24
  my0           = 24
  mz0           = 24
  boundary_x    = 1
  boundary_y    = 1
  boundary_z    = 0
<output>
  interval      = {last}
  istep         = 2
  dump          = 1
  report        = 2
<problem>
  L0            = 4.0
  r_L0          = 1.0
  t0            = 5.0
  beta_twist    = 0.3
"""


def test_fortran_uniform_deck_matches_oracle(tmp_path):
    """user_uniform's twisted flux tube: the driver against the oracle, the
    userOutput datasets (var1 = rho, var2 = |E x B|_perp / (r |B|^2)) and the
    timing report of every second step."""
    last = 8
    out, stdout = _run("tristan_ff_uniform", UNIFORM_DECK.format(last=last), tmp_path, "uniform")
    p = problems.uniform(24, L0=4.0, r_L0=1.0, beta_twist=0.3, t0=5.0)
    o = Oracle(p.n, boundary=p.boundary, cc=p.cc)
    problems.load(o, p)
    o.preloop()
    o.run(0, last + 1)
    got = _read_dump(out / f"flds.{last:05d}.bin", 24)
    err = field_errors(lambda f: got[f], o.field)
    print("fortran uniform", err)
    assert max(err.values()) <= 1e-12
    # var2 from the raw components at the nodes, in default REAL (FP32)
    sl = np.s_[2:-2:2, 2:-2:2, 2:-2:2]
    c = {f: got[f][sl].astype(np.float32) for f in FIELDS6}
    a = -c["bz"] * c["ex"] + c["bx"] * c["ez"]
    b = c["bz"] * c["ey"] - c["by"] * c["ez"]
    k, j, i = np.meshgrid(np.arange(12), np.arange(12), np.arange(12), indexing="ij")
    d1 = (2 * i).astype(np.float32) - np.float32(12.0)
    d2 = (2 * j).astype(np.float32) - np.float32(12.0)
    var2 = np.sqrt(a * a + b * b) / np.sqrt(d1 * d1 + d2 * d2 + np.float32(1e-7)) / (
        c["bx"] * c["bx"] + c["by"] * c["by"] + c["bz"] * c["bz"])
    got2 = h5_read(out / "flds.tot.00001", "var2", np.float32, tmp_path)
    assert np.allclose(got2, var2.ravel(), rtol=1e-6, atol=0)
    # makeReport (tristanmainloop.F90:108-151) after steps 0, 2, 4, ...
    assert " [ROUTINE]          [TIME, ms]      [MIN  /  MAX, ms]      [FRACTION, %]" in stdout
    assert re.search(r"^ Timestep: 4\.+\[DONE\] $", stdout, re.M)  # column 72 is a blank
    assert not re.search(r"^ Timestep: 3\.", stdout, re.M)
    for row in (" Full_step: ", "   fld_solver: ", "     exchange: ", "   usr_funcs: ", "   output_step: "):
        assert re.search("^" + re.escape(row), stdout, re.M), row


def test_fortran_shipped_loop_deck_stops_as_the_reference(tmp_path):
    """inputs/input.loop lacks t0 and beta_twist, which user_loop reads
    without defaults: the reference stops there, and so does the driver."""
    deck = """<node_configuration>
  sizex         = 4
  sizey         = 4
  sizez         = 7
<time>
  last           = 1000
<grid>
  mx0           = 256
  my0           = 256
  mz0           = 259
  boundary_x    = 1
  boundary_y    = 1
  boundary_z    = 0
<output>
  start         = 0
  interval      = 100
  istep         = 4
<problem>
  a_L0          = 3.0
  h_L0          = 2.0
  L0            = 20.0
  r_L0          = 0.5
  v_twist       = 0.3
"""
    path = os.path.join(BUILD, "tristan_ff_loop")
    if not os.path.exists(path):
        pytest.skip("tristan_ff_loop not built")
    (tmp_path / "input").write_text(deck)
    r = subprocess.run([path, "-i", str(tmp_path / "input"), "-o", str(tmp_path / "out")], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-1000:])
    assert r.returncode != 0
    assert "ERROR variable `t0` not found in <problem>" in r.stdout + r.stderr
