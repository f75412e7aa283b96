#!/usr/bin/env python3
"""Runs the named workloads end to end on one GPU and reports throughput and
the FP32-bulk path's error against the FP64 path on the same steps.

  emwave64   vacuum plane wave, 64^3 periodic, CC = 0.5, 200 steps; both
             paths are also compared with the closed-form SSP-RK3 Yee
             solution;
  alfven256  B0 = z + seeded Alfvenic dB/B0 = 0.05 (k <= 8), 256^3 periodic,
             CC = 0.45, 500 steps;
  loop384    twisted loop (user_loop field; the footpoint driver on z <= 0
             and a conducting top, seeded total angle), 384^3, x/y periodic,
             z open, 1000 steps;
  psr512     aligned rotator (user_psr form), R* = 20, R_LC = 80, absorbing
             layer 40 cells (fixed mode), 512^3, CC = 0.5, two periods
             (about 2011 steps);
  scan48     48 loop / plane-wave instances with seeded edges in [64, 192],
             300 steps, on the batched path, and the same instances run one
             after another on the single-meshblock path.

Per workload: cell-updates/s of each path; for FP32-bulk, for FP32 (the
reference's own precision), for the compensated state (FP32 pairs with
two-sum stage accumulation, on the per-procedure kernels) and for FP64
started from B perturbed by one part in 1e12 (how much the workload itself
amplifies roundoff), against FP64: the energy history's relative
error (on E_EM and on E_EM - E_EM(t=0)), the normwise max and L2 relative
error of every component after the last step, and div_h B of each run at
t = 0 and after the last step with its change over the run (over every
interior cell and over the cells no hook or absorbing layer writes, for
which the scheme conserves it).  One JSON object per workload is printed,
and appended to --out if given.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import problems  # noqa: E402
import pytff  # noqa: E402

F6 = ("ex", "ey", "ez", "bx", "by", "bz")
NG = problems.NG


def log(msg: str):
    print(f"[configs] {msg}", flush=True)


def _interior(a):
    return a[NG:-NG, NG:-NG, NG:-NG]


def run_single(p, precision: str, steps: int, every: int, keep_fields: bool):
    """Energy history every `every` steps, wall time of the stepping, final
    interior fields (FP64 on the host)."""
    mb = pytff.Meshblock(p.n, boundary=p.boundary, cc=p.cc, precision=precision,
                         absorb_buff=p.absorb_buff, absorb_fixed=p.absorb_fixed)
    problems.load(mb, p)
    mb.preloop()
    hist = [mb.reduce()]
    b0 = [mb.download(f) for f in ("bx", "by", "bz")]
    bmax = max(float(np.abs(_interior(a)).max()) for a in b0)
    d0 = problems.div_h_b(*b0)
    del b0
    mb.run(0, 2)  # autotune and graph capture, timed apart
    mb.sync()
    done, wall = 2, 0.0
    t_last = time.time()
    while done < steps:
        k = min(every - done % every if done % every else every, steps - done)
        mb.sync()
        t = time.perf_counter()
        mb.run(done, k)
        mb.sync()
        wall += time.perf_counter() - t
        done += k
        if done % every == 0 or done == steps:
            hist.append(mb.reduce())
        if time.time() - t_last > 20:
            log(f"  {p.name} {precision}: step {done}/{steps}")
            t_last = time.time()
    d1 = problems.div_h_b(*(mb.download(f) for f in ("bx", "by", "bz")))
    out = {"hist": np.array(hist), "cups": p.cells * (steps - 2) / wall, "ms_per_step": wall * 1e3 / (steps - 2),
           "tuning": mb.get_tuning(), "divb": problems.divb_report(p, d0, d1, bmax)}
    del d0, d1
    if keep_fields:
        out["fields"] = {f: _interior(mb.download(f)).copy() for f in F6}
    mb.close()
    return out


def perturbation_energy(p) -> float:
    """1/2 sum over the interior of E^2 + (B - B0)^2 at t = 0."""
    f = p.fields
    tot = 0.0
    for c in "xyz":
        tot += float(np.sum(_interior(f["e" + c]) ** 2))
        tot += float(np.sum((_interior(f["b" + c]) - _interior(f["b0" + c])) ** 2))
    return 0.5 * tot


def compare(p, r64, r32):
    """Energy errors: on E_EM relative to E_EM, and on the energy change
    E_EM(t) - E_EM(0) relative to the larger of its FP64 maximum and the
    perturbation energy at t = 0 (a plane wave conserves energy, so its
    energy change alone is roundoff)."""
    e64 = r64["hist"][:, 0] + r64["hist"][:, 1]
    e32 = r32["hist"][:, 0] + r32["hist"][:, 1]
    e_rel = float(np.max(np.abs(e32 - e64) / np.abs(e64)))
    d64, d32 = e64 - e64[0], e32 - e32[0]
    scale = max(float(np.max(np.abs(d64))), perturbation_energy(p), 1e-12 * abs(float(e64[0])))
    de_rel = float(np.max(np.abs(d32 - d64)) / scale) if scale > 0 else 0.0
    res = {"energy_rel_err_max": e_rel, "delta_energy_rel_err_max": de_rel,
           "energy_drift_fp64": float((e64[-1] - e64[0]) / e64[0])}
    if "fields" in r64 and "fields" in r32:
        comp = {}
        for f in F6:
            a, b = r64["fields"][f], r32["fields"][f]
            ok = np.isfinite(a) & np.isfinite(b)
            ref = float(np.max(np.abs(a[ok]))) if ok.any() else 0.0
            d = np.abs(a[ok] - b[ok])
            comp[f] = {"max_rel": float(d.max() / ref) if ref > 0 else float(d.max(initial=0.0)),
                       "l2_rel": float(np.sqrt(np.sum(d * d) / max(np.sum(a[ok] ** 2), 1e-300))),
                       "nonfinite": int((~np.isfinite(a)).sum() + (~np.isfinite(b)).sum())}
        res["fields"] = comp
    res["divb"] = {"fp64": r64["divb"], "other": r32["divb"]}
    return res


def closed_form_err(p, res, amp, mode, steps):
    """Max error of the final B_z perturbation against the closed-form
    SSP-RK3 Yee plane wave, relative to the wave amplitude."""
    ref = problems.emwave_closed_form(p, amp, mode, steps)
    return float(np.max(np.abs((res["fields"]["bz"] - 1.0) - ref))) / amp


def single_workload(name, p, steps, every, extra=None):
    """FP64, FP32-bulk, FP32 (the reference's own precision) and the
    compensated state (FP32 pairs, two-sum stage accumulation; per-procedure
    kernels) on the same steps; each is compared with FP64."""
    log(f"{name}: {p.n[0]}^3, {steps} steps")
    runs = {}
    for prec in ("fp64", "fp32_bulk", "fp32", "fp32_comp"):
        runs[prec] = run_single(p, prec, steps, every, True)
        log(f"  {prec} {runs[prec]['cups']:.4g} cell-updates/s")
    # FP64 again from B perturbed by one part in 1e12: how far roundoff-size
    # differences grow on this workload by themselves
    pp = dataclasses.replace(p, fields=dict(p.fields))
    pp.fields["bz"] = p.fields["bz"] * (1.0 + 1e-12)
    rpert = run_single(pp, "fp64", steps, every, True)
    log(f"  fp64 with B perturbed by 1e-12: done")
    r64 = runs["fp64"]
    out = {"workload": name, "n": list(p.n), "steps": steps, "boundary": list(p.boundary), "cc": p.cc,
           "cells": p.cells, "g": p.g}
    word = {"fp64": 8, "fp32_bulk": 4, "fp32": 4, "fp32_comp": 8}
    for prec, r in runs.items():
        out[prec] = {"cups": r["cups"], "ms_per_step": r["ms_per_step"], "tuning": r["tuning"],
                     "roofline_frac_canonical": r["cups"] * word[prec] * (48 + 9 * p.g) / 8.0e12}
    out["fp32_bulk_over_fp64"] = runs["fp32_bulk"]["cups"] / r64["cups"]
    out["fp32_bulk_vs_fp64"] = compare(p, r64, runs["fp32_bulk"])
    out["fp32_vs_fp64"] = compare(p, r64, runs["fp32"])
    out["fp32_comp_vs_fp64"] = compare(p, r64, runs["fp32_comp"])
    out["fp64_perturbed_1e-12_vs_fp64"] = compare(p, r64, rpert)
    h = {prec: (r["hist"][:, 0] + r["hist"][:, 1]).tolist() for prec, r in runs.items()}
    h["fp64_perturbed_1e-12"] = (rpert["hist"][:, 0] + rpert["hist"][:, 1]).tolist()
    h["sample_every"] = every
    out["energy_history"] = h
    if extra:
        out.update(extra(p, runs))
    return out


def scan_workload(count, steps, seed):
    probs = problems.scan(count=count, nmin=64, nmax=192, seed=seed, steps=steps)
    count = len(probs)
    cells = sum(p.cells for p in probs)
    log(f"scan{count}: {count} instances, {cells} cells, {steps} steps")
    out = {"workload": f"scan{count}", "instances": count, "cells": cells, "steps": steps,
           "edges": [p.n[0] for p in probs], "kinds": [p.name.rstrip("0123456789") for p in probs]}
    hists = {}
    # FP64 again from B perturbed by one part in 1e12 (batched): how far
    # roundoff-size differences grow on each instance by themselves
    pert = []
    for p in probs:
        pp = dataclasses.replace(p, fields=dict(p.fields))
        pp.fields["bz"] = p.fields["bz"] * (1.0 + 1e-12)
        pert.append(pp)
    for prec in ("fp64", "fp32_bulk", "fp64_perturbed"):
        b = pytff.Batch(problems.batch_specs(probs), precision=prec.replace("_perturbed", ""))
        for q, p in enumerate(pert if prec == "fp64_perturbed" else probs):
            problems.load(b[q], p)
            b[q].preloop()
        h0 = [b[q].reduce() for q in range(count)]
        d0, bmax = [], []
        for q in range(count if prec != "fp64_perturbed" else 0):
            bq = [b[q].download(f) for f in ("bx", "by", "bz")]
            bmax.append(max(float(np.abs(_interior(a)).max()) for a in bq))
            d0.append(problems.div_h_b(*bq))
        b.run(0, 2)
        t = time.perf_counter()
        b.run(2, steps - 2)
        for q in range(count):
            b[q].sync()
        wall = time.perf_counter() - t
        hists[prec] = (np.array(h0), np.array([b[q].reduce() for q in range(count)]))
        if prec == "fp64_perturbed":
            b.close()
            continue
        reps = [problems.divb_report(p, d0[q], problems.div_h_b(*(b[q].download(f) for f in ("bx", "by", "bz"))),
                                     bmax[q]) for q, p in enumerate(probs)]
        del d0
        out[prec + "_divb"] = {
            "change_max_free_rel_bmax_worst": max(r["change_max_free_rel_bmax"] for r in reps),
            "max_t0_worst": max(r["max_t0"] for r in reps), "max_end_worst": max(r["max_end"] for r in reps)}
        out[prec] = {"batch_cups": cells * (steps - 2) / wall, "batch_ms_per_step": wall * 1e3 / (steps - 2),
                     "tuning": b.get_tuning()}
        b.close()
        log(f"  batch {prec}: {out[prec]['batch_cups']:.4g} cell-updates/s")
        # the same instances one after another on the single-meshblock path
        wall1 = 0.0
        for p in probs:
            mb = pytff.Meshblock(p.n, boundary=p.boundary, cc=p.cc, precision=prec)
            problems.load(mb, p)
            mb.preloop()
            mb.run(0, 2)
            mb.sync()
            t = time.perf_counter()
            mb.run(2, steps - 2)
            mb.sync()
            wall1 += time.perf_counter() - t
            mb.close()
        out[prec]["sequential_cups"] = cells * (steps - 2) / wall1
        log(f"  one by one {prec}: {out[prec]['sequential_cups']:.4g} cell-updates/s")
    (a0, a1), (b0, b1) = hists["fp64"], hists["fp32_bulk"]
    e64, e32 = a1[:, 0] + a1[:, 1], b1[:, 0] + b1[:, 1]
    d64, d32 = e64 - (a0[:, 0] + a0[:, 1]), e32 - (b0[:, 0] + b0[:, 1])
    scale = np.maximum(np.abs(d64), np.array([perturbation_energy(p) for p in probs]))
    scale = np.maximum(scale, 1e-12 * np.abs(e64))
    out["energy_rel_err_max"] = float(np.max(np.abs(e32 - e64) / np.abs(e64)))
    ep = hists["fp64_perturbed"][1][:, 0] + hists["fp64_perturbed"][1][:, 1]
    out["energy_rel_err_per_instance"] = [float(x) for x in np.abs(e32 - e64) / np.abs(e64)]
    out["perturbed_fp64_energy_rel_err_per_instance"] = [float(x) for x in np.abs(ep - e64) / np.abs(e64)]
    out["perturbed_fp64_energy_rel_err_max"] = float(np.max(np.abs(ep - e64) / np.abs(e64)))
    out["delta_energy_rel_err_max"] = float(np.max(np.abs(d32 - d64) / np.maximum(scale, 1e-300)))
    out["batch_over_sequential_fp32_bulk"] = out["fp32_bulk"]["batch_cups"] / out["fp32_bulk"]["sequential_cups"]
    out["batch_over_sequential_fp64"] = out["fp64"]["batch_cups"] / out["fp64"]["sequential_cups"]
    # canonical HBM fraction: the algorithmic bytes of every instance's
    # cell-updates (w (48 + 9 g) per cell-update) over the batch's wall time
    for prec, word in (("fp64", 8), ("fp32_bulk", 4)):
        bytes_step = sum(p.cells * word * (48 + 9 * p.g) for p in probs)
        wall_step = out[prec]["batch_ms_per_step"] / 1e3
        out[prec]["roofline_frac_canonical"] = bytes_step / wall_step / 8.0e12
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("workloads", nargs="*", default=["emwave64", "alfven256", "loop384", "psr512", "scan48"])
    ap.add_argument("--out", default="")
    ap.add_argument("--scale", type=float, default=1.0,
                    help="multiply every step count by this (short rehearsals)")
    a = ap.parse_args(argv)

    def steps(n):
        return max(4, int(round(n * a.scale)))

    results = []
    for w in a.workloads:
        t0 = time.time()
        if w == "emwave64":
            amp, mode = 1e-2, (1, 1)
            p = problems.emwave(64, amp=amp, mode=mode, cc=0.5)
            ns = steps(200)
            res = single_workload(w, p, ns, 10, lambda p, runs: {
                "closed_form_err_rel_amp": {prec: closed_form_err(p, r, amp, mode, ns)
                                            for prec, r in runs.items()}})
        elif w == "alfven256":
            res = single_workload(w, problems.alfven(256, seed=1, kmax=8), steps(500), 50)
        elif w == "loop384":
            ns = steps(1000)
            res = single_workload(w, problems.loop(384, steps=ns, driver=problems.LOOP_DRIVEN), ns, 100)
        elif w == "psr512":
            p = problems.psr(512)
            ns = steps(p.steps)
            res = single_workload(w, p, ns, 100)
        elif w == "scan48":
            res = scan_workload(48, steps(300), 11)
        else:
            raise SystemExit(f"unknown workload {w}")
        res["wall_s"] = time.time() - t0
        print(json.dumps(res), flush=True)
        results.append(res)
        if a.out:
            with open(a.out, "a") as fh:
                fh.write(json.dumps(res) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
