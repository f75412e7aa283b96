"""Build libtff.so (HIP kernels + C ABI) for gfx950, in-tree.

    python tristan-ff_amd/build.py            # libtff.so
    python tristan-ff_amd/build.py --fortran  # also the Fortran host (tristan_ff)

The library is written to tristan-ff_amd/_build/libtff.so, next to the
sources, so that it travels with the tree to the machine that runs it.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
OUT = os.path.join(PKG, "_build")
LIB = os.path.join(OUT, "libtff.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: no fused multiply-add, so the FP64/FP32 paths round
# exactly where the reference's expressions round.  Packed FP32 arithmetic
# is disabled for every device function.
HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-I", os.path.join(PKG, "csrc"),
]


def build_lib(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    src = os.path.join(PKG, "csrc", "tff.hip")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, src, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


PROBE = os.path.join(OUT, "stream_probe")


def build_probe(verbose: bool = False) -> str:
    """stream_probe: the streaming bandwidth of the stage kernel's mix of HBM
    streams with no arithmetic (the ceiling its roofline fraction is read
    against)."""
    os.makedirs(OUT, exist_ok=True)
    src = os.path.join(PKG, "csrc", "stream_probe.hip")
    tmp = PROBE + ".tmp"
    cmd = [HIPCC, "-O3", f"--offload-arch={ARCH}", "-std=c++17",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", src, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, PROBE)
    return PROBE


HDF5_ROOT = os.environ.get("HDF5_ROOT", "/opt/conda")
IO_LIB = os.path.join(OUT, "libtff_io.so")


def build_io(verbose: bool = False) -> str:
    """libtff_io.so: the output files of writeOutput through the serial HDF5
    C API (host C, linked against libtff.so and libhdf5).  Skipped, with a
    message, when no HDF5 headers are found."""
    inc = os.path.join(HDF5_ROOT, "include")
    lib = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"hdf5.h not found under {inc}: libtff_io.so not built", flush=True)
        return ""
    src = os.path.join(PKG, "csrc", "tff_io.c")
    tmp = IO_LIB + ".tmp"
    # RUNPATH (not RPATH): the HDF5 directory applies to this library's own
    # dependencies only, never to the HIP runtime's C++ library
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "include"), "-I", inc, src, "-o", tmp,
           "-L", OUT, "-ltff", "-L", lib, "-lhdf5", "-lm",
           "-Wl,--enable-new-dtags", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{lib}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, IO_LIB)
    return IO_LIB


def build_fortran(verbose: bool = False) -> str:
    """Fortran host (amdflang): deck parser, main loop and user files that
    call libtff.so through ISO_C_BINDING."""
    fdir = os.path.join(PKG, "fortran")
    mk = os.path.join(fdir, "Makefile")
    if not os.path.exists(mk):
        return ""
    subprocess.run(["make", "-s", "-C", fdir], check=True)
    return os.path.join(OUT, "tristan_ff")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--fortran", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    print(build_lib(a.verbose))
    print(build_io(a.verbose))
    print(build_probe(a.verbose))
    if a.fortran and shutil.which("amdflang"):
        print(build_fortran(a.verbose))
    return 0


if __name__ == "__main__":
    sys.exit(main())
