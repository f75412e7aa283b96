#!/usr/bin/env python3
"""HBM traffic per launch of the fused stage kernel from rocprofv3 counter
runs, calibrated by a plain copy kernel of known size.

Input: four ``--pmc`` runs of profile_run.py, one counter each (FETCH_SIZE
and WRITE_SIZE of a ``step`` run, the same two of a ``copy`` run), each a
rocprofv3 ``--output-format csv`` directory, plus the ``--info`` JSON files
those runs wrote.  The counters read in KiB per dispatch; the copy kernel
moves a known number of bytes each way, so each counter's factor is true
bytes / counter bytes for the copy, and the stage kernel's traffic is its
counters times those factors, averaged over the timed launches.

Output: ``<out>.json`` (the figure bench.py reports as roofline.traffic,
with the kernel-source hash it was measured on) and ``<out>.md`` (the
table and the commands).
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def per_dispatch(run_dir: str, counter: str, want: str, skip: str = "fix") -> list[tuple[str, float]]:
    """(kernel name, counter value in bytes) of every dispatch whose kernel
    name contains `want` and not `skip`, in dispatch order."""
    paths = glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True)
    if not paths:
        raise FileNotFoundError(f"no counter_collection.csv under {run_dir}")
    acc: dict = collections.defaultdict(float)
    names: dict = {}
    for path in paths:
        with open(path) as fh:
            for r in csv.DictReader(fh):
                k = r["Kernel_Name"]
                if r["Counter_Name"] != counter or want not in k or (skip and skip in k):
                    continue
                d = int(r["Dispatch_Id"])
                acc[d] += float(r["Counter_Value"]) * 1024.0  # KiB -> bytes
                names[d] = k
    return [(names[d], acc[d]) for d in sorted(acc)]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--step-fetch", required=True)
    ap.add_argument("--step-write", required=True)
    ap.add_argument("--copy-fetch", required=True)
    ap.add_argument("--copy-write", required=True)
    ap.add_argument("--step-info", required=True, help="JSON written by profile_run.py step --info")
    ap.add_argument("--copy-info", required=True, help="JSON written by profile_run.py copy --info")
    ap.add_argument("--out", required=True, help="output path without extension")
    ap.add_argument("--commands", default="", help="file holding the commands that made the runs")
    a = ap.parse_args(argv)
    si = json.load(open(a.step_info))
    ci = json.load(open(a.copy_info))
    copy_true = float(ci["bytes_each_way_per_launch"])
    factors = {}
    for c, run in (("FETCH_SIZE", a.copy_fetch), ("WRITE_SIZE", a.copy_write)):
        v = [x for _, x in per_dispatch(run, c, "k_copy3")]
        tail = v[-2 * ci["calls"]:]
        factors[c] = copy_true / (sum(tail) / len(tail))
    launches = 3 * si["timed_steps"]
    rows = {}
    kname = None
    for c, run in (("FETCH_SIZE", a.step_fetch), ("WRITE_SIZE", a.step_write)):
        d = per_dispatch(run, c, "k_stage")
        if len(d) < launches:
            raise RuntimeError(f"{run}: {len(d)} stage dispatches, {launches} timed launches expected")
        d = d[-launches:]
        kname = d[0][0]
        rows[c] = [[x for _, x in d[s::3]] for s in range(3)]  # per RK3 stage
    fetch = [sum(v) / len(v) * factors["FETCH_SIZE"] for v in rows["FETCH_SIZE"]]
    write = [sum(v) / len(v) * factors["WRITE_SIZE"] for v in rows["WRITE_SIZE"]]
    rd, wr = sum(fetch) / 3.0, sum(write) / 3.0
    alg = si["alg_bytes_per_launch"]
    out = {"kernel": kname.split("(")[0], "workload": si["workload"], "precision": si["precision"],
           "tuning": si["tuning"], "bytes_per_launch": rd + wr, "read_bytes_per_launch": rd,
           "write_bytes_per_launch": wr, "alg_bytes_per_launch": alg, "traffic_over_alg": (rd + wr) / alg,
           "fetch_factor": factors["FETCH_SIZE"], "write_factor": factors["WRITE_SIZE"],
           "per_stage_read": fetch, "per_stage_write": write, "source_hash": si["source_hash"]}
    with open(a.out + ".json", "w") as fh:
        json.dump(out, fh)
        fh.write("\n")
    cmds = open(a.commands).read().strip() if a.commands else "(see the runs' directories)"
    mb = lambda x: f"{x / 1e6:,.1f} MB"
    lines = [f"# HBM traffic of the stage kernel ({si['precision']}, {si['workload']})", "",
             f"Kernel `{out['kernel']}`, tuning {si['tuning']}, kernel sources {si['source_hash']}.",
             f"{si['timed_steps']} timed steps ({launches} stage launches) after the warm-up.", "",
             "Commands (one counter per run, from the repository root):", "", "```", cmds, "```", "",
             "## Calibration", "",
             f"`k_copy3` (tff_init_rk_step) moves {mb(copy_true)} each way per launch.", "",
             "| counter | factor (true bytes / counter bytes) |", "|---|---|",
             f"| FETCH_SIZE | {factors['FETCH_SIZE']:.3f} |", f"| WRITE_SIZE | {factors['WRITE_SIZE']:.3f} |", "",
             "## Stage kernel, per launch", "",
             "| RK3 stage | read (calibrated) | written (calibrated) |", "|---|---|---|"]
    for s in range(3):
        lines.append(f"| {s + 1} | {mb(fetch[s])} | {mb(write[s])} |")
    lines += ["", f"- **Measured traffic:** {mb(rd)} read + {mb(wr)} written = **{mb(rd + wr)}** per launch.",
              f"- **Algorithmic bytes:** {mb(alg)} per launch; measured / algorithmic = "
              f"**{(rd + wr) / alg:.3f}**.", ""]
    with open(a.out + ".md", "w") as fh:
        fh.write("\n".join(lines))
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
