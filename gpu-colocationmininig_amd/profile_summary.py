#!/usr/bin/env python3
"""Summaries of rocprofv3 output for the miner's kernels.

  kernels  <run_results.db | kernel_trace.csv>  [--top N] [--templates]
      Markdown table of per-kernel calls and durations (microseconds); the
      instantiations of a kernel template are summed unless --templates.
  traffic  --fetch F.csv --write W.csv --cfetch CF.csv --cwrite CW.csv --kernel NAME [--copy-bytes B]
           [--workload-key TEXT]
      HBM bytes per launch of NAME from two `rocprofv3 --pmc FETCH_SIZE` /
      `--pmc WRITE_SIZE` runs (KiB per dispatch), each scaled by what the same
      counter reads for a plain device copy of B bytes (bench.py --copy-probe,
      __amd_rocclr_copyBuffer): scale = B / counter.  Prints JSON.
"""
import argparse
import collections
import csv
import json
import os
import sqlite3


def _kernel_rows(path):
    """(name, duration_ns) for every dispatch in a trace database or CSV."""
    if path.endswith(".db"):
        db = sqlite3.connect(path)
        return [(n, float(d)) for n, d in db.execute("select name, duration from kernels")]
    rows = []
    with open(path) as fh:
        for r in csv.DictReader(fh):
            if "Kernel_Name" in r and "Start_Timestamp" in r:
                rows.append((r["Kernel_Name"], float(r["End_Timestamp"]) - float(r["Start_Timestamp"])))
    return rows


def kernels(path, top, templates=False):
    agg = collections.defaultdict(list)
    for n, d in _kernel_rows(path):
        base = n.split("(")[0].replace("void ", "").strip() or n
        if not templates:
            base = base.split("<")[0]
        agg[base].append(d)
    tot = sum(sum(v) for v in agg.values())
    out = ["| kernel | calls | total us | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for n, v in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:top]:
        out.append(f"| `{n}` | {len(v)} | {sum(v) / 1e3:.1f} | {sum(v) / len(v) / 1e3:.1f} | {min(v) / 1e3:.1f} | "
                   f"{max(v) / 1e3:.1f} | {100 * sum(v) / tot:.1f} |")
    out.append(f"\nTotal kernel time {tot / 1e6:.1f} ms over {sum(len(v) for v in agg.values())} dispatches.")
    return "\n".join(out)


def _counter_per_dispatch(path, counter, kernel):
    """Mean of `counter` (summed over its dimensions) per dispatch of kernel."""
    per = collections.defaultdict(float)
    with open(path) as fh:
        for r in csv.DictReader(fh):
            if r["Counter_Name"] != counter or kernel not in r["Kernel_Name"]:
                continue
            per[r["Dispatch_Id"]] += float(r["Counter_Value"])
    if not per:
        raise SystemExit(f"no {counter} samples for {kernel} in {path}")
    return sum(per.values()) / len(per), len(per)


def traffic(a):
    f, nd = _counter_per_dispatch(a.fetch, "FETCH_SIZE", a.kernel)
    w, _ = _counter_per_dispatch(a.write, "WRITE_SIZE", a.kernel)
    cf, _ = _counter_per_dispatch(a.cfetch, "FETCH_SIZE", "__amd_rocclr_copyBuffer")
    cw, _ = _counter_per_dispatch(a.cwrite, "WRITE_SIZE", "__amd_rocclr_copyBuffer")
    fs = a.copy_bytes / (cf * 1024.0)
    ws = a.copy_bytes / (cw * 1024.0)
    return {"kernel": a.kernel, "dispatches": nd, "fetch_kib_per_dispatch_raw": f, "write_kib_per_dispatch_raw": w,
            "fetch_scale": fs, "write_scale": ws, "hbm_bytes_per_launch": f * 1024.0 * fs + w * 1024.0 * ws,
            "calibration": f"plain copy of {a.copy_bytes} B: FETCH_SIZE {cf:.1f} KiB, WRITE_SIZE {cw:.1f} KiB "
                           f"per dispatch", "sources": [os.path.basename(p) for p in (a.fetch, a.write)],
            "workload_key": a.workload_key}


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("path")
    k.add_argument("--top", type=int, default=25)
    k.add_argument("--templates", action="store_true", help="one line per template instantiation")
    t = sub.add_parser("traffic")
    for n in ("fetch", "write", "cfetch", "cwrite", "kernel"):
        t.add_argument(f"--{n}", required=True)
    t.add_argument("--copy-bytes", type=float, default=256 * 1024 * 1024)
    t.add_argument("--workload-key", default=None, help="the workload description bench.py prints in config.workload")
    a = ap.parse_args()
    if a.cmd == "kernels":
        print(kernels(a.path, a.top, a.templates))
    else:
        print(json.dumps(traffic(a), indent=1))


if __name__ == "__main__":
    main()
