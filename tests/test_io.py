"""libtff_io.so (the output files of writeOutput, write_output.F90:27-500)
through its C ABI, read back with the HDF5 tools (h5dump): dataset names,
types, shapes and values of params.NNNNN, flds.tot.NNNNN and domain.NNNNN,
and the XDMF wrapper's text.  Runs without a GPU."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tristan-ff_amd", "_build", "libtff_io.so")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"

pytestmark = pytest.mark.skipif(not os.path.exists(LIB) or not os.path.exists(H5DUMP),
                                reason="libtff_io.so or h5dump not present (no HDF5 here)")


def _lib():
    lib = ctypes.CDLL(LIB)
    lib.tff_io_write_params.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_float)]
    lib.tff_io_write_fields.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_void_p)]
    lib.tff_io_write_domain.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.tff_io_last_error.restype = ctypes.c_char_p
    return lib


def _header(path):
    return subprocess.run([H5DUMP, "-H", path], check=True, capture_output=True, text=True).stdout


def _values(path, name, dtype, tmp):
    out = os.path.join(tmp, name.replace(":", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path], check=True, capture_output=True)
    return np.fromfile(out, dtype=dtype)


def test_params_file(tmp_path):
    lib = _lib()
    params = [("cpu:sizex", 1, 4, 0.0), ("grd:mx0", 1, 256, 0.0), ("alg:c", 2, 0, 0.45),
              ("prb:psr_angle", 2, 0, 30.0), ("out:flag", 3, 1, 0.0)]
    names = b"".join(n.encode() + b"\0" for n, *_ in params)
    types = (ctypes.c_int * 5)(*[t for _, t, _, _ in params])
    ivals = (ctypes.c_int * 5)(*[i for _, _, i, _ in params])
    rvals = (ctypes.c_float * 5)(*[r for *_, r in params])
    d = str(tmp_path).encode()
    assert lib.tff_io_write_params(d, 7, 1400, 5, names, types, ivals, rvals) == 0, lib.tff_io_last_error()
    path = os.path.join(tmp_path, "params.00007")
    hdr = _header(path)
    for n, t, i, r in params:
        block = hdr[hdr.index(f'DATASET "{n}"'):]
        block = block[:block.index("}") + 1]
        assert ("H5T_IEEE_F32LE" if t == 2 else "H5T_STD_I32LE") in block, n
        assert "( 1 )" in hdr[hdr.index(f'DATASET "{n}"'):][:300], n
        v = _values(path, n, np.float32 if t == 2 else np.int32, str(tmp_path))
        assert v.size == 1 and v[0] == (np.float32(r) if t == 2 else i), n
    assert _values(path, "timestep", np.int32, str(tmp_path))[0] == 1400
    # an unknown parameter type is an error, as in writeParams_hdf5
    types[0] = 4
    assert lib.tff_io_write_params(d, 8, 0, 5, names, types, ivals, rvals) != 0


def test_fields_file_and_xdmf(tmp_path):
    lib = _lib()
    ni, nj, nk, istep = 5, 4, 3, 2
    rng = np.random.default_rng(3)
    data = [rng.standard_normal((nk, nj, ni)).astype(np.float32) for _ in range(9)]
    ptrs = (ctypes.c_void_p * 9)(*[a.ctypes.data for a in data])
    dims = (ctypes.c_int * 3)(ni, nj, nk)
    d = str(tmp_path).encode()
    assert lib.tff_io_write_fields(d, 12, istep, dims, ptrs) == 0, lib.tff_io_last_error()
    path = os.path.join(tmp_path, "flds.tot.00012")
    hdr = _header(path)
    names = ["ex", "ey", "ez", "bx", "by", "bz", "var1", "var2", "var3", "xx", "yy", "zz"]
    assert re.findall(r'DATASET "(\w+)"', hdr) == sorted(names)
    for n in names:
        block = hdr[hdr.index(f'DATASET "{n}"'):][:400]
        assert "H5T_IEEE_F32LE" in block and f"( {nk}, {nj}, {ni} )" in block, n
    for q, n in enumerate(names[:9]):
        assert np.array_equal(_values(path, n, np.float32, str(tmp_path)), data[q].ravel()), n
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    for n, c in zip(("xx", "yy", "zz"), (i, j, k)):
        assert np.array_equal(_values(path, n, np.float32, str(tmp_path)), (c * istep).astype(np.float32).ravel())
    xdmf = open(path + ".xdmf").read().splitlines()
    assert xdmf[0] == '<?xml version="1.0" ?>'
    assert xdmf[5] == '      <Topology TopologyType="3DCoRectMesh" Dimensions="         3         4         5"/>'
    assert xdmf.count('          flds.tot.00012:/var2') == 1
    assert sum(1 for s in xdmf if s.startswith('      <Attribute Name="')) == 12
    assert xdmf[-1] == "</Xdmf>"


def test_domain_file(tmp_path):
    lib = _lib()
    blocks = (ctypes.c_int * 12)(0, 0, 0, 64, 64, 32, 0, 0, 32, 64, 64, 32)
    d = str(tmp_path).encode()
    assert lib.tff_io_write_domain(d, 0, 2, blocks) == 0, lib.tff_io_last_error()
    path = os.path.join(tmp_path, "domain.00000")
    want = {"x0": [0, 0], "y0": [0, 0], "z0": [0, 32], "sx": [64, 64], "sy": [64, 64], "sz": [32, 32]}
    for n, v in want.items():
        assert list(_values(path, n, np.int32, str(tmp_path))) == v, n
