#!/usr/bin/env python3
"""Writes tests/golden/ref_field_solver.npz: input and output vectors of the
reference's own field solver (src/algorithms/field_solver.F90, compiled from
the reference sources into oracle/_ref by oracle/Makefile), so that the CPU
oracle stays pinned to it where oracle/_ref cannot be built.

Each case: seeded FP64 fields E, B, B0 (ghosts included, Fortran layout;
stored once per boundary mix, rounded for the FP32 cases) on a
6 x 5 x 4 box with CC = 0.45 and absorb_buff = 2; then initRKstep and three
(rk3Step(c1, c2, c3), cleanEpar) stages with the SSP-RK3 coefficients, and
the resulting E, B.  A few cells get |E| > |B| so that the limiter runs.
Run from the repository root after `make -C oracle`.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from helpers import FIELDS6, random_fields  # noqa: E402
from oracle.oracle import RefSolver, rk_coefficients  # noqa: E402

N = (6, 5, 4)
CASES = [("f64", (1, 1, 1)), ("f64", (0, 1, 0)), ("f64", (2, 2, 2)),
         ("f32", (1, 1, 1)), ("f32", (0, 1, 0)), ("f32", (2, 2, 2))]


def main():
    out = {}
    shape = (N[2] + 4, N[1] + 4, N[0] + 4)
    for name, boundary in CASES:
        dt = np.float64 if name == "f64" else np.float32
        if not RefSolver.available(dt):
            raise SystemExit("oracle/_ref is not built")
        tag = f"{name}_{''.join(map(str, boundary))}"
        rng = np.random.default_rng(20261203 + 100 * boundary[0] + 10 * boundary[1] + boundary[2])
        fl = random_fields(shape, rng)  # FP64; the FP32 cases round these
        fl["ex"][4, 4, 4] = 40.0
        fl["ey"][5, 3, 3] = -25.0
        r = RefSolver(N, boundary=boundary, cc=0.45, absorb_buff=2, dtype=dt)
        for f, a in fl.items():
            r.put(f, a.astype(dt))
            if name == "f64":
                out[f"in_{''.join(map(str, boundary))}/{f}"] = a
        r.init_rk_step()
        co = rk_coefficients(dt)
        for s in range(3):
            r.rk3_step(*co[3 * s:3 * s + 3])
            r.clean_epar()
        for f in FIELDS6:
            out[f"{tag}/out/{f}"] = r.get(f)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_field_solver.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
