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
  counters --trace T.csv --pmc P1.csv [P2.csv ...] --fetch F.csv --write W.csv --cfetch CF.csv --cwrite CW.csv
           [--match k_count,k_write] [--min-ms 1]
      Markdown table, one line per dispatch of the matching kernels lasting
      at least min-ms, of the counter runs of one deterministic program (the
      n-th dispatch of a kernel is the same work in every run; dispatch ids
      may differ): duration, TA / TD busy and TD stalls
      (per unit and cycle, 256 units), L2 hit rate, VALU and vector-read
      instructions, VALU busy (4 cycles per instruction, 1024 SIMDs) and HBM
      bytes (traffic's copy-calibrated scales).  Cycles per XCD =
      GRBM_GUI_ACTIVE / 8.
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


def _ordinal_keys(pairs):
    """(kernel name, n) for each (dispatch id, kernel name), n counting that
    kernel's dispatches in id order."""
    seen = collections.Counter()
    out = {}
    for did, name in sorted(pairs, key=lambda x: int(x[0])):
        out[did] = (name, seen[name])
        seen[name] += 1
    return out


def _per_dispatch(path):
    """{(kernel name, n): {counter: value summed over dimensions}} for the
    n-th dispatch of each kernel."""
    vals, names = {}, {}
    with open(path) as fh:
        for r in csv.DictReader(fh):
            names[r["Dispatch_Id"]] = r["Kernel_Name"]
            vals.setdefault(r["Dispatch_Id"], collections.defaultdict(float))[r["Counter_Name"]] += float(r["Counter_Value"])
    keys = _ordinal_keys(names.items())
    return {keys[d]: v for d, v in vals.items()}


def counters(a):
    dur = {}
    with open(a.trace) as fh:
        disp = [(r["Dispatch_Id"], r["Kernel_Name"], (float(r["End_Timestamp"]) - float(r["Start_Timestamp"])) / 1e6)
                for r in csv.DictReader(fh)]
    keys = _ordinal_keys([(d, n) for d, n, _ in disp])
    for d, n, ms in disp:
        dur[keys[d]] = (n, ms)
    merged = collections.defaultdict(dict)
    for p in a.pmc + [a.fetch, a.write]:
        for key, cs in _per_dispatch(p).items():
            merged[key].update(cs)
    cf, _ = _counter_per_dispatch(a.cfetch, "FETCH_SIZE", "__amd_rocclr_copyBuffer")
    cw, _ = _counter_per_dispatch(a.cwrite, "WRITE_SIZE", "__amd_rocclr_copyBuffer")
    fs, ws = a.copy_bytes / (cf * 1024.0), a.copy_bytes / (cw * 1024.0)
    pats = a.match.split(",")
    rows = ["| kernel | ms | TA busy | TD busy | TD stalled on L1 | L2 hit | VALU instr | VMEM read instr | VALU busy | "
            "HBM read GB | HBM write GB |", "|---|---|---|---|---|---|---|---|---|---|---|"]
    order = {keys[d]: int(d) for d, _, _ in disp}
    for did in sorted(dur, key=lambda x: order[x]):
        name, ms = dur[did]
        short = name.split("(")[0].replace("void ", "").strip()
        if ms < a.min_ms or not any(short.startswith(p) for p in pats) or did not in merged:
            continue
        c = merged[did]
        cyc = c.get("GRBM_GUI_ACTIVE", 0.0) / 8.0
        if not cyc:
            continue
        hit, miss = c.get("TCC_HIT_sum", 0.0), c.get("TCC_MISS_sum", 0.0)
        rows.append(f"| `{short}` | {ms:.1f} | {c.get('TA_TA_BUSY_sum', 0) / (256 * cyc):.0%} | "
                    f"{c.get('TD_TD_BUSY_sum', 0) / (256 * cyc):.0%} | {c.get('TD_TC_STALL_sum', 0) / (256 * cyc):.0%} | "
                    f"{hit / (hit + miss) if hit + miss else 0:.2f} | {c.get('SQ_INSTS_VALU', 0):.3g} | "
                    f"{c.get('SQ_INSTS_VMEM_RD', 0):.3g} | {c.get('SQ_INSTS_VALU', 0) * 4 / (1024 * cyc):.0%} | "
                    f"{c.get('FETCH_SIZE', 0) * 1024 * fs / 1e9:.1f} | {c.get('WRITE_SIZE', 0) * 1024 * ws / 1e9:.1f} |")
    return "\n".join(rows)


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
    c = sub.add_parser("counters")
    c.add_argument("--trace", required=True)
    c.add_argument("--pmc", nargs="+", required=True)
    for n in ("fetch", "write", "cfetch", "cwrite"):
        c.add_argument(f"--{n}", required=True)
    c.add_argument("--match", default="k_count,k_write")
    c.add_argument("--min-ms", type=float, default=1.0)
    c.add_argument("--copy-bytes", type=float, default=256 * 1024 * 1024)
    a = ap.parse_args()
    if a.cmd == "kernels":
        print(kernels(a.path, a.top, a.templates))
    elif a.cmd == "counters":
        print(counters(a))
    else:
        print(json.dumps(traffic(a), indent=1))


if __name__ == "__main__":
    main()
