#!/usr/bin/env python3
"""Benchmark of the force-free step on one MI355X.

Headline workload: uniform B0 = z^ plus a seeded Alfvenic perturbation
(|dB|/B0 = 0.05, 1 <= |k| <= 8), 256^3 periodic Yee grid, CC = 0.45; each
timed step is one full SSP-RK3 timestep (three stages, each with the
force-free current, the update, the E.B = 0 projection and the ghost fill),
inputs resident in HBM.  A second workload is measured the same way and
reported in the same line: the aligned rotator (user_psr form) at 512^3, a
dipole of radius 20 cells, light cylinder at 80 cells, a 40-cell fixed
absorbing layer on every side and the conducting-sphere hook every stage.
Initial conditions are generated on the device.  Both paths run the same
steps from the same initial state:

* fp32_bulk -- FP32 storage and arithmetic (fused multiply-adds, the
  scale-exact stage update), FP64 redo near magnetic nulls, FP64
  reductions (the headline ``value``);
* fp64      -- FP64 throughout (parity-checked against the FP64 oracle).

A third workload, the driven twisted loop at 384^3 (x/y periodic, z open,
the footpoint driver on the floor and a conducting top), is measured the
same way.

The FP32-bulk energy is compared with the FP64 path's energy after the run.
``roofline`` is for the dominant kernel (the stage kernel of the FP32-bulk
path): algorithmic bytes per launch over its average launch time, the timed
region's HIP events (on the library's own stream) over its stage launches;
``traffic`` is the
calibrated FETCH_SIZE + WRITE_SIZE per launch stored under profiles/ by
traffic.py, reported only when it was measured on the kernel sources being
run (their hash is compared).

``cpu_baseline`` is the reference's own field_solver.F90 (oracle/_ref) on
the host: one process on one core, and the aggregate of eight concurrent
single-core processes (the reference's parallelism is MPI ranks, one core
each), each on a bounded sample of the same workload.  It runs before the
GPU is touched.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "tristan-ff_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
METRIC = "cell-updates/s + %HBM roofline: FP32-bulk vs FP64 at ≤1e-5 rel. energy err"


def algorithmic_bytes_per_cell_update(word: int, g: int) -> int:
    """w * (48 + 9 g): stage 1 reads U^n and writes U^1 (12 words), stages 2
    and 3 read U^n and U^(s-1) and write U^s (18 words each), plus 3 words of
    curl B0 per stage when B0 is non-uniform."""
    return word * (48 + 9 * g)


def run_path(precision: str, prob, steps: int, warmup: int, timing_steps: int, tuning=None):
    import pytff
    import problems
    mb = pytff.Meshblock(prob.n, boundary=prob.boundary, cc=prob.cc, precision=precision,
                         absorb_buff=prob.absorb_buff, absorb_fixed=prob.absorb_fixed)
    if tuning is not None:
        mb.set_tuning(*tuning)
    problems.load(mb, prob)
    mb.preloop()
    e_start = mb.reduce()
    # div_h B at t = 0, on the host (outside every timed region)
    b3 = [mb.download(f) for f in ("bx", "by", "bz")]
    bmax = max(float(abs(a[2:-2, 2:-2, 2:-2]).max()) for a in b3)
    d0 = problems.div_h_b(*b3)
    del b3
    mb.run(0, warmup)
    mb.sync()
    t0 = time.perf_counter()
    mb.event(0)
    mb.run(warmup, steps)
    mb.event(1)
    mb.sync()
    ms = (time.perf_counter() - t0) * 1e3
    ms_events = mb.elapsed_ms(0, 1)
    e_end = mb.reduce()
    divb = problems.divb_report(prob, d0, problems.div_h_b(*(mb.download(f) for f in ("bx", "by", "bz"))), bmax)
    del d0
    # the stage kernel's launches in the timed region, read from HIP events
    # on the library's stream around that region: its event time over the
    # 3 * steps stage launches.  Whatever else the region launches (the
    # near-null fix-up scans, ghost passes and hooks) counts as stage-kernel
    # time, so this bounds the kernel's average duration from above; it
    # agrees with rocprofv3's kernel trace of the same command within the
    # fix-up scans' few microseconds.
    region_launch_s = ms_events / 1e3 / (3 * steps)
    # for reference: the same steps continued eagerly, each stage launch
    # bracketed by events (eager launches run a few per cent slower than the
    # graph-replayed ones)
    mb.kernel_timing(True)
    for s in range(timing_steps):
        mb.step(warmup + steps + s)
    ktot, klaunch = mb.stage_kernel_time()
    mb.kernel_timing(False)
    cells = prob.cells
    word = 8 if precision == "fp64" else 4
    b_alg = algorithmic_bytes_per_cell_update(word, prob.g)
    cups = cells * steps / (ms / 1e3)
    eager_launch_s = ktot / 1e3 / max(klaunch, 1)
    # the three stage launches move 12 + 18 + 18 words per cell: 16 on average
    per_launch_bytes = cells * word * (48 + 9 * prob.g) / 3.0
    achieved = per_launch_bytes / region_launch_s
    tun = mb.get_tuning()
    mb.close()
    return {
        "precision": precision,
        "value": cups,
        "ms_per_step": ms / steps,
        "ms_per_step_events": ms_events / steps,
        "tuning": tun,
        "b_alg_per_cell_update": b_alg,
        "roofline_frac_canonical": cups * b_alg / HBM_PEAK,
        "stage_kernel": {"avg_ms": region_launch_s * 1e3, "launches": 3 * steps,
                         "alg_bytes_per_launch": per_launch_bytes,
                         "achieved_GBps": achieved / 1e9,
                         "eager_event_avg_ms": eager_launch_s * 1e3, "eager_launches": klaunch},
        "energy_end": e_end[0] + e_end[1],
        "energy_start": e_start[0] + e_start[1],
        "divb": divb,
    }


def cpu_sample(n: int, budget_s: float) -> dict:
    """Timesteps of the reference's field_solver.F90 (oracle/_ref, FP32 as
    shipped, amdflang -O3) on the Alfvenic workload at n^3 until about
    `budget_s` seconds have been timed; step loop only."""
    import numpy as np
    from oracle.oracle import RefSolver
    import problems
    if not RefSolver.available(np.float32):
        raise FileNotFoundError("oracle/_ref not built")
    p = problems.alfven(n, seed=1, kmax=8)
    r = RefSolver(p.n, boundary=p.boundary, cc=p.cc, dtype=np.float32)
    for f, a in p.fields.items():
        r.put(f, a.astype(np.float32))
    r.exchange()
    t = time.perf_counter()
    r.run(1)
    el = time.perf_counter() - t
    nsteps = 1
    more = min(59, int(budget_s / max(el, 1e-9)) - 1)
    if more > 0:
        t = time.perf_counter()
        r.run(more)
        el += time.perf_counter() - t
        nsteps += more
    return {"cells": p.cells, "steps": nsteps, "seconds": el}


def _cpu_workers(n: int, budget_s: float, cpus: list) -> list:
    """One bench.py --cpu-worker child per CPU, started together."""
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", str(n), str(budget_s),
                               str(c)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for c in cpus]
    out = []
    for pr in procs:
        so, se = pr.communicate(timeout=budget_s * 6 + 120)
        if pr.returncode != 0:
            raise RuntimeError(f"cpu worker failed ({pr.returncode}): {se.strip()[-300:]}")
        out.append(json.loads(so.strip().splitlines()[-1]))
    return out


def cpu_baseline(n: int, budget_s: float, ncores: int = 8) -> dict:
    """The reference's solver on the host: one single-core process alone,
    then `ncores` single-core processes at once (each pinned to its own
    CPU), summed.  Reported, never required."""
    res = {"value": None, "unit": "cell-updates/s", "cores": 1, "kind": "reference"}
    try:
        cpus = sorted(os.sched_getaffinity(0))
        one = _cpu_workers(n, budget_s, cpus[:1])[0]
        res["value"] = one["cells"] * one["steps"] / one["seconds"]
        res["sample"] = (f"{one['steps']} timestep(s) of the Alfvenic workload at {n}^3, reference "
                         f"field_solver.F90 built FP32 as shipped (amdflang -O3), single rank, serial "
                         f"periodic ghost fill, {one['seconds']:.1f} s on one core")
        use = cpus[:ncores]
        many = _cpu_workers(n, budget_s, use)
        res["aggregate"] = {"value": sum(w["cells"] * w["steps"] / w["seconds"] for w in many),
                            "unit": "cell-updates/s", "cores": len(use),
                            "sample": f"{len(use)} concurrent single-core processes, each "
                                      f"{min(w['steps'] for w in many)}-{max(w['steps'] for w in many)} "
                                      f"timesteps of the same workload, rates summed"}
    except Exception as e:  # the baseline is reported, never required
        res["sample"] = f"unavailable: {e}"
    return res


def load_traffic(workload: str, source_hash: str):
    """The stored calibrated traffic of this workload's stage kernel, if it
    was measured on the same kernel sources."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None, f"no profiles/traffic_{workload}.json"
    with open(path) as fh:
        tr = json.load(fh)
    if tr.get("source_hash") != source_hash:
        return None, f"profiles/traffic_{workload}.json was measured on other kernel sources"
    return tr, f"profiles/traffic_{workload}.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; copy-kernel calibrated)"


# Stage-kernel shapes the benchmarks pin, (tile index, z-chunk) per path:
# the fastest shape of a sweep of tiles 0-5 with z-chunks 16, 32 and 64 on
# each workload (two interleaved rounds, 40 graph-replayed steps each), so
# that the line and a profile of the same command time the same kernel.
# Tile indices: 0 32x4, 1 32x8, 2 64x4, 3 64x8, 4 32x16, 5 64x16.
PINNED = {
    "alfven": {"fp64": (3, 64), "fp32_bulk": (3, 64)},
    "psr": {"fp64": (1, 32), "fp32_bulk": (4, 32)},
    "loop": {"fp64": (1, 64), "fp32_bulk": (4, 16)},
}


def measure_workload(name: str, prob, steps: int, warmup: int, timing_steps: int, tuning) -> dict:
    """Both paths on one workload, FP32-bulk's energy against FP64's;
    `tuning` maps each path to its pinned (tile, z-chunk), or is None (the
    autotuner)."""
    res = {}
    for prec in ("fp64", "fp32_bulk"):
        res[prec] = run_path(prec, prob, steps, warmup, timing_steps, tuning[prec] if tuning else None)
        print(f"[bench] {name} {prec}: {res[prec]['value']:.4g} cell-updates/s, "
              f"{res[prec]['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    e64, e32 = res["fp64"]["energy_end"], res["fp32_bulk"]["energy_end"]
    de64 = e64 - res["fp64"]["energy_start"]
    de32 = e32 - res["fp32_bulk"]["energy_start"]
    res["energy_rel_err"] = abs(e32 - e64) / abs(e64)
    res["delta_energy_rel_err"] = abs(de32 - de64) / max(abs(de64), 1e-300)
    return res


def roofline_of(head: dict, workload: str) -> dict:
    import pytff
    roof = {
        "bound": "hbm",
        "achieved": head["stage_kernel"]["achieved_GBps"],
        "peak": HBM_PEAK / 1e9,
        "unit": "GB/s",
        "frac": head["stage_kernel"]["achieved_GBps"] / (HBM_PEAK / 1e9),
        "traffic": None,
        "kernel": "k_stage<PrecF32Bulk, %dx%d tile> (fused stage kernel, kchunk %d)" % (
            head["tuning"]["tx"], head["tuning"]["ty"], head["tuning"]["kchunk"]),
        "avg_launch_ms": head["stage_kernel"]["avg_ms"],
        "avg_launch_source": "timed region's HIP events / stage launches (includes the fix-up scans)",
        "eager_event_avg_ms": head["stage_kernel"]["eager_event_avg_ms"],
        "alg_bytes_per_launch": head["stage_kernel"]["alg_bytes_per_launch"],
    }
    tr, src = load_traffic(workload + "_fp32_bulk", pytff.kernel_source_hash())
    if tr and tr.get("tuning") == head["tuning"]:
        roof["traffic"] = tr["bytes_per_launch"]
    elif tr:
        src = f"{src}: measured with tuning {tr.get('tuning')}, not {head['tuning']}"
    roof["traffic_source"] = src
    return roof


def summary(res: dict, head: dict, workload: str) -> dict:
    return {"value": head["value"], "ms_per_step": head["ms_per_step"], "roofline": roofline_of(head, workload),
            "roofline_frac_canonical": head["roofline_frac_canonical"],
            "fp64": {"value": res["fp64"]["value"], "ms_per_step": res["fp64"]["ms_per_step"],
                     "roofline_frac_canonical": res["fp64"]["roofline_frac_canonical"],
                     "stage_kernel": res["fp64"]["stage_kernel"]},
            "fp32_bulk_over_fp64": head["value"] / res["fp64"]["value"],
            "energy_rel_err_fp32_bulk_vs_fp64": res["energy_rel_err"],
            "delta_energy_rel_err_fp32_bulk_vs_fp64": res["delta_energy_rel_err"],
            "divb": {"fp64": res["fp64"]["divb"], "fp32_bulk": head["divb"],
                     "note": "div_h B at cell centres: max at t = 0 and at the end, and its change over the "
                             "run over the cells no hook or absorbing layer writes (_free), where the "
                             "scheme conserves it"}}


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if argv[:1] == ["--cpu-worker"]:
        n, budget, cpu = int(argv[1]), float(argv[2]), int(argv[3])
        os.sched_setaffinity(0, {cpu})
        print(json.dumps(cpu_sample(n, budget)))
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--timing-steps", type=int, default=20)
    ap.add_argument("--psr-n", type=int, default=512, help="rotator box (0: leave the rotator out)")
    ap.add_argument("--psr-steps", type=int, default=200)
    ap.add_argument("--loop-n", type=int, default=384, help="driven-loop box (0: leave the loop out)")
    ap.add_argument("--loop-steps", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU work per baseline process")
    ap.add_argument("--tuning", default="pinned", choices=("pinned", "auto"),
                    help="pinned: the stage kernel's tile and z-chunk that the autotuner picks on "
                         "each workload in most runs (PINNED), so that the line and a profile of the "
                         "same command time the same kernel; auto: the autotuner")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: the force-free step runs on one GPU only (asked for --gpus {a.gpus}); "
              "sharding the box over several GPUs is designed but not built or measured.",
              file=sys.stderr)
        return 2
    # the host baseline first: its worker processes start before this
    # process touches the GPU.  The reference's loops run the slowest (z)
    # index innermost, about 3 us per cell-update on one core: a bounded
    # sample at 128^3 (same generator and per-cell work; its smaller planes
    # favour the CPU's caches)
    cpu = None if a.no_cpu_baseline else cpu_baseline(min(a.n, 128), a.cpu_seconds)
    if cpu:
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)
    import problems
    t0 = time.time()
    prob = problems.alfven(a.n, seed=1, kmax=8, host=False)
    workload = f"alfven{a.n}"
    pin = a.tuning == "pinned"
    res = measure_workload(workload, prob, a.steps, a.warmup, a.timing_steps, PINNED["alfven"] if pin else None)
    print(f"[bench] {workload} done in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    head = res["fp32_bulk"]
    main_sum = summary(res, head, workload)
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{workload}: B0=z + Alfvenic dB/B0=0.05 (k<=8, seed 1), "
                               f"{a.n}^3 periodic, CC=0.45, SSP-RK3 force-free step",
                   "cells": prob.cells, "g": prob.g, "parallelism": "single GPU"},
        "roofline": main_sum["roofline"],
    }
    line.update({k: v for k, v in main_sum.items() if k not in ("value", "ms_per_step", "roofline")})
    if a.psr_n:
        t0 = time.time()
        psr = problems.psr(a.psr_n, host=False)
        wl = f"psr{a.psr_n}"
        rp = measure_workload(wl, psr, a.psr_steps, 10, 5, PINNED["psr"] if pin else None)
        print(f"[bench] {wl} done in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
        ps = summary(rp, rp["fp32_bulk"], wl)
        ps["steps"] = a.psr_steps
        ps["config"] = {"workload": f"{wl}: aligned rotator, dipole radius 20 cells, light cylinder 80 cells, "
                                    f"40-cell fixed absorbing layer, {a.psr_n}^3, CC=0.5",
                        "cells": psr.cells, "g": psr.g}
        line.setdefault("workloads", {})[wl] = ps
    if a.loop_n:
        t0 = time.time()
        loop = problems.loop(a.loop_n, steps=1000, driver=problems.LOOP_DRIVEN, host=False)
        wl = f"loop{a.loop_n}"
        rl = measure_workload(wl, loop, a.loop_steps, 10, 5, PINNED["loop"] if pin else None)
        print(f"[bench] {wl} done in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
        ls = summary(rl, rl["fp32_bulk"], wl)
        ls["steps"] = a.loop_steps
        ls["tuning"] = {"fp32_bulk": rl["fp32_bulk"]["tuning"], "fp64": rl["fp64"]["tuning"]}
        ls["config"] = {"workload": f"{wl}: twisted flux loop (user_loop field, footpoint driver on z <= 0, "
                                    f"conducting top), {a.loop_n}^3, x/y periodic, z open, CC=0.45",
                        "cells": loop.cells, "g": loop.g}
        line.setdefault("workloads", {})[wl] = ls
    line["cpu_baseline"] = cpu
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
