"""The C ABI library loads and exports every entry point include/tff.h
declares (no GPU needed: only symbol resolution, no compute calls)."""
import ctypes
import os

import pytest

import pytff


def test_header_declares_the_reference_procedures():
    names = set(pytff.header_functions())
    for n in ("tff_init_rk_step", "tff_rk3_step", "tff_clean_epar", "tff_exchange_fields",
              "tff_user_bc", "tff_step", "tff_run", "tff_create", "tff_destroy", "tff_upload",
              "tff_download", "tff_reduce", "tff_set_bc", "tff_last_error",
              "tff_output_dims", "tff_output_fields"):
        assert n in names, n


def test_library_exports_every_declared_symbol():
    if not os.path.exists(pytff.LIB_PATH):
        pytest.skip("libtff.so not built")
    lib = ctypes.CDLL(pytff.LIB_PATH)
    missing = [n for n in pytff.header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_config_struct_matches_header_layout():
    # nx, ny, nz, boundary[3], absorb_buff, absorb_fixed, precision, device (10 ints), cc
    assert ctypes.sizeof(pytff.Config) == 10 * 4 + 8
    assert pytff.Config.cc.offset == 40


def test_no_cpu_fallback_without_the_library():
    """With the library missing, the Python host raises instead of
    computing anything on the CPU."""
    import subprocess
    import sys
    env = dict(os.environ, TFF_LIBRARY="/nonexistent/libtff.so")
    code = ("import pytff\n"
            "try:\n"
            "    pytff.Meshblock(8)\n"
            "except pytff.TffError as e:\n"
            "    print('raised:', e)\n")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120,
                       cwd=os.path.dirname(pytff.__file__))
    assert "raised:" in r.stdout and "not built" in r.stdout, (r.stdout, r.stderr)


def test_no_cpu_fallback_without_a_gpu():
    """On a machine without a GPU, creating a meshblock fails with an error
    from the C ABI rather than running elsewhere."""
    if not os.path.exists(pytff.LIB_PATH):
        pytest.skip("libtff.so not built")
    try:
        g = pytff.Meshblock(8)
    except pytff.TffError:
        return
    g.close()
    pytest.skip("a GPU is present: the CPU-only behaviour is not observable here")
