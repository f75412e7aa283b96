#!/usr/bin/env python3
"""Short, fixed workloads for hardware-counter and kernel-trace runs.

  step  -- a few fused timesteps of a benchmark workload at one precision:
           alfven (B0 = z + seeded dB/B0 = 0.05, k <= 8, periodic), psr
           (the aligned rotator: dipole of radius 20 cells, light cylinder
           at 80 cells, 40-cell fixed absorbing layer on every side) or
           loop (the driven twisted loop: x/y periodic, z open);
  copy  -- tff_init_rk_step on the same grid: two launches of a plain copy
           kernel (k_copy3), E -> En and B -> Bn over the allocated box,
           3 * nalloc words each way per launch -- a kernel of known
           traffic, against which FETCH_SIZE / WRITE_SIZE are calibrated;
  output -- tff_output_fields (k_output: writeFields' node interpolation
           at every istep-th node) on the same workload after two steps.

Each mode prints what it ran (sizes, tile, bytes) so that a counter summary
can be matched to it, and with --info writes the same as JSON (traffic.py
reads it).  Initial conditions are generated on the device.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import problems  # noqa: E402
import pytff  # noqa: E402


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=("step", "copy", "output"))
    ap.add_argument("--precision", default="fp32_bulk", choices=("fp64", "fp32", "fp32_bulk"))
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--tile", default="", help="tile:kchunk (default: autotuned)")
    ap.add_argument("--istep", type=int, default=2, help="output mode: node stride")
    ap.add_argument("--workload", default="alfven", choices=("alfven", "psr", "loop"))
    ap.add_argument("--info", default="", help="write what ran as JSON to this file")
    a = ap.parse_args(argv)
    if a.workload == "alfven":
        p = problems.alfven(a.n, seed=1, kmax=8, host=False)
    elif a.workload == "loop":
        p = problems.loop(a.n, steps=1000, driver=problems.LOOP_DRIVEN, host=False)
    else:
        p = problems.psr(a.n, host=False)
    mb = pytff.Meshblock(p.n, boundary=p.boundary, cc=p.cc, precision=a.precision, absorb_buff=p.absorb_buff,
                         absorb_fixed=p.absorb_fixed)
    problems.load(mb, p)
    mb.preloop()
    word = 8 if a.precision == "fp64" else 4

    def info(**kw):
        if a.info:
            kw.update(workload=f"{a.workload}{a.n}", precision=a.precision, source_hash=pytff.kernel_source_hash())
            with open(a.info, "w") as fh:
                json.dump(kw, fh)
    if a.mode == "copy":
        mb.initRKstep()  # allocates En, Bn
        mb.sync()
        for _ in range(a.steps):
            mb.initRKstep()
        mb.sync()
        nalloc = mb.nalloc
        print(f"copy: {a.steps} calls, 2 k_copy3 launches each (E, then B), nalloc {nalloc}, "
              f"{3 * nalloc * word} bytes read and {3 * nalloc * word} written per launch", flush=True)
        info(mode="copy", calls=a.steps, bytes_each_way_per_launch=3 * nalloc * word)
        mb.close()
        return 0
    if a.mode == "output":
        mb.run(0, 2)
        mb.output_fields(a.istep)  # the state is now clean; later calls time the output alone
        t0 = time.perf_counter()
        for _ in range(a.steps):
            out = mb.output_fields(a.istep)
        ms = (time.perf_counter() - t0) * 1e3 / max(a.steps, 1)
        nout = out["ex"].size
        print(f"output: {a.precision} {a.n}^3, istep {a.istep}, {nout} nodes x 6 datasets, "
              f"{6 * p.cells * word} bytes of fields read (each cell once) and {6 * nout * 4} written "
              f"per k_output launch; {ms:.3f} ms per call including the copies to the host", flush=True)
        mb.close()
        return 0
    if a.tile:
        t, kc = a.tile.split(":")
        mb.set_tuning(int(t), int(kc))
    mb.run(0, 2)
    mb.sync()
    t0 = time.perf_counter()
    mb.run(2, a.steps)
    mb.sync()
    ms = (time.perf_counter() - t0) * 1e3 / max(a.steps, 1)
    tun = mb.get_tuning()
    alg = p.cells * word * (48 + 9 * p.g) / 3.0
    print(f"step: {a.workload} {a.precision} {a.n}^3, tuning {tun}, {a.steps} steps, "
          f"{ms:.3f} ms/step, {alg:.0f} algorithmic bytes per stage launch", flush=True)
    info(mode="step", timed_steps=a.steps, tuning=tun, alg_bytes_per_launch=alg, ms_per_step=ms)
    mb.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
