"""The Fortran driver's deck reading (read_input.F90:56-385 semantics) on the
CPU: every error below stops the driver before it touches a GPU."""
import os
import subprocess

import pytest

BUILD = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tristan-ff_amd", "_build")
EXE = os.path.join(BUILD, "tristan_ff_emwave")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (no amdflang)")


def _run(tmp_path, deck):
    (tmp_path / "input").write_text(deck)
    r = subprocess.run([EXE, "-i", str(tmp_path / "input"), "-o", str(tmp_path / "out")], capture_output=True,
                       text=True, timeout=60)
    return r.returncode, r.stdout + r.stderr


def test_missing_key_without_default_stops(tmp_path):
    rc, out = _run(tmp_path, "<node_configuration>\n  sizex = 1\n<grid>\n  mx0 = 16\n")
    assert rc != 0
    assert "ERROR variable `sizey` not found in <node_configuration>\n-- no default value provided" in out


def test_unconvertible_value_stops(tmp_path):
    rc, out = _run(tmp_path, "<node_configuration>\n  sizex = 1\n  sizey = 1\n  sizez = abc   # z\n")
    assert rc != 0
    assert "ERROR variable `sizez`=abc from <node_configuration> not converted" in out


def test_key_is_the_first_list_directed_token(tmp_path):
    """`mz0=16` (no blank) is one token for a list-directed read, so the key
    is not found; a key after the next block header is not in the block;
    blank records are skipped; comments after '#' are not part of values."""
    deck = ("<node_configuration>\n  sizex = 1\n\n  sizey = 1\n  sizez = 1 # cpus\n"
            "<grid>\n  mx0 = 16\n  my0 = 16\n  mz0=16\n")
    rc, out = _run(tmp_path, deck)
    assert rc != 0 and "ERROR variable `mz0` not found in <grid>" in out
    deck = ("<node_configuration>\n  sizex = 1\n  sizey = 1\n  sizez = 1\n<grid>\n  mx0 = 16\n  my0 = 16\n"
            "<time>\n  mz0 = 16\n")
    rc, out = _run(tmp_path, deck)
    assert rc != 0 and "ERROR variable `mz0` not found in <grid>" in out


def test_grid_must_divide_by_the_cpu_layout(tmp_path):
    deck = ("<node_configuration>\n  sizex = 3\n  sizey = 1\n  sizez = 1\n<grid>\n  mx0 = 16\n  my0 = 16\n"
            "  mz0 = 16\n")
    rc, out = _run(tmp_path, deck)
    assert rc != 0 and "ERROR: grid size is not evenly divisible by the number of cores" in out


def test_istep_one_is_rejected(tmp_path):
    deck = ("<node_configuration>\n  sizex = 1\n  sizey = 1\n  sizez = 1\n<grid>\n  mx0 = 16\n  my0 = 16\n"
            "  mz0 = 16\n<output>\n  istep = 1\n")
    rc, out = _run(tmp_path, deck)
    assert rc != 0 and "istep = 1" in out
