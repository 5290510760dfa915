#!/usr/bin/env python3
"""Benchmark of the colocation-mining hot path on one MI355X.

A step is one full mining job on the workload held in HBM: grid binning
(colo_bin) and the Apriori level loop (filter, neighbour-cell joins, PI
numerators, instance tables) up to the workload's largest pattern size, at
every prevalence threshold of the workload in one colo_mine_sweep call
(theta-independent results are computed once per step and reused).

The default workload is c5: 50M instances, 60 features with Zipf(1.2)
frequencies, 60% of the points in 2,000 Gaussian hotspots (sigma log-uniform
in 100-2000 on a 1e5 square, 12 features per hotspot), 20 planted patterns of
size 3-6, r for ~64 mean neighbours, patterns up to size 6, thresholds 0.1,
0.2, 0.3 and 0.5.  The data is generated in memory with a fixed seed
(float64 coordinates converted to float32, the reference's coordinate type)
before the timed region.  c1/c2/c3 are available with --workload.

Prints one JSON line.  value = neighbour pairs per second of the level-2
join (pairs counted by the level-2 count pass / wall time of the level-2
count + PI + scan + write); mining_s = wall time of a step.  The timed
schedule is the learned one (persistent launches whose work queue is ordered
and partitioned, and whose table arenas are pre-sized, by the trained cost
model in gpu-colocationmininig_amd/models/cost_mlp.bin) when that model is
present, else the persistent one; the persistent and plain (per-candidate)
schedules also run on the same data and every report must be identical.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-colocationmininig_amd"))

METRIC = "end-to-end mining s + neighbor pairs/s; learned-sched vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def describe(p):
    th = ", ".join(f"{t:g}" for t in p["thetas"])
    return (f"{p['name']}: {p['N']} instances, {p['F']} features, r={p['r']:g}, theta in {{{th}}}, "
            f"patterns up to size {p['max_k']}, seed {p['seed']}")


def prepare(args):
    import datagen
    import numpy as np
    t0 = time.perf_counter()
    labels, x, y, S, p = datagen.generate(args.workload, N=args.N)
    fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    log(f"generated {describe(p)} in {time.perf_counter() - t0:.1f} s")
    return x32, y32, fc, np.array([0, 0, S, S], np.float32), p


def cpu_baseline(x, y, fc, bounds, p, budget_s, threads):
    """The oracle's level-2 joins (count + write) of the same workload on
    host cores, over as many candidate pairs as fit the time budget."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.set_threads(threads)
    o = oracle.Oracle.from_arrays(x, y, fc, bounds)
    o.bin(p["r"])
    F = o.F
    pairs = [(a, b) for a in range(F) for b in range(a + 1, F)]
    # spread the sample over frequent and rare features
    pairs.sort(key=lambda ab: (ab[0] + ab[1]) % 7)
    rows = 0
    t_tot = 0.0
    used = 0
    for a, b in pairs:
        t0 = time.perf_counter()
        n, slots, _, _ = o.join2_count(a, b)
        o.join2_write(a, b, slots)
        t_tot += time.perf_counter() - t0
        rows += n
        used += 1
        if t_tot > budget_s:
            break
    return {"value": rows / t_tot if t_tot else 0.0, "unit": "neighbor pairs/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle level-2 joins (count + write) of {used} of {len(pairs)} feature pairs of "
                      f"{p['name']} ({rows} pairs, {t_tot:.1f} s, OpenMP {threads} threads)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5")
    ap.add_argument("--N", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--model", default=os.path.join(ROOT, "gpu-colocationmininig_amd", "models", "cost_mlp.bin"))
    ap.add_argument("--no-model", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--tune-cache", default=os.path.join(os.path.expanduser("~"), ".cache", "colo_tune.txt"))
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy 3 times (traffic-counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(64 * 1024 * 1024, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4, "copies": 3}))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (multi-GPU sharding is a design "
                                   "note only)", "n_gpus": args.gpus}))
        return 2

    x, y, fc, bounds, p = prepare(args)

    import colo
    import torch

    ctx = colo.Context(x, y, fc, bounds, p["r"])
    torch.cuda.synchronize()

    model = None
    if not args.no_model and os.path.exists(args.model):
        model = colo.CostModel.load(args.model)
    main_sched = colo.COLO_SCHED_LEARNED if model is not None else colo.COLO_SCHED_PERSISTENT
    sched_name = {colo.COLO_SCHED_LEARNED: "learned", colo.COLO_SCHED_PERSISTENT: "persistent",
                  colo.COLO_SCHED_PLAIN: "plain"}
    thetas = p["thetas"]

    tune = ctx.tune
    if not args.no_tune:
        t0 = time.perf_counter()
        os.makedirs(os.path.dirname(args.tune_cache), exist_ok=True)
        tune = ctx.autotune(min(thetas), max_k=p["max_k"], sched=main_sched, model=model, cache=args.tune_cache)
        log(f"autotuned launch shapes in {time.perf_counter() - t0:.1f} s: {tune}")

    bin_ms = []

    def step(sched):
        t0 = time.perf_counter()
        ctx.rebin(p["r"])
        torch.cuda.synchronize()
        bin_ms.append((time.perf_counter() - t0) * 1000.0)
        out = ctx.mine_sweep(thetas, max_k=p["max_k"], sched=sched, hash_tables=False,
                             model=model if sched == colo.COLO_SCHED_LEARNED else None)
        return [o[0] for o in out], [o[1] for o in out]

    for i in range(args.warmup):
        step(main_sched)
        log(f"warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        rep, sts = step(main_sched)
        stats.extend(sts)
        log(f"step {i + 1}/{args.steps}: " + ", ".join(f"theta {t:g}: {s['total_ms']:.0f} ms / {s['patterns']} patterns"
                                                      for t, s in zip(thetas, sts)))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps

    pairs = sum(s["l2_pairs"] for s in stats)
    l2_s = sum(s["l2_join_ms"] for s in stats) / 1000.0
    value = pairs / l2_s if l2_s else 0.0

    out = {"metric": METRIC, "value": value, "unit": "neighbor pairs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32",
           "data": f"synthetic: seeded generator ({p['name']}), generated in memory",
           "config": {"workload": describe(p), "instances": p["N"], "features": p["F"], "r": p["r"],
                      "thetas": thetas, "max_k": p["max_k"], "schedule": sched_name[main_sched],
                      "launch_shapes": tune}}
    out["mining_s"] = {sched_name[main_sched]: ms_per_step / 1000.0}
    if model is not None:
        n_m = sum(s["model_n"] for s in stats)
        out["cost_model"] = {"mean_abs_log1p_rows_err": sum(s["model_abs_err"] for s in stats) / n_m if n_m else None,
                             "overflow_fallbacks": sum(s["overflows"] for s in stats),
                             "arenas": sum(s["arenas"] for s in stats), "predictions": n_m}
    out["phases_ms_per_step"] = {
        "bin": sum(bin_ms[args.warmup:args.warmup + args.steps]) / args.steps,
        **{k: sum(s[k] for s in stats) / args.steps for k in ("filter_ms", "model_ms", "count_ms", "scan_ms",
                                                               "write_ms", "host_ms", "total_ms")}}
    out["patterns"] = {f"{t:g}": s["patterns"] for t, s in zip(thetas, sts)}
    out["levels"] = {f"{t:g}": s["levels"] for t, s in zip(thetas, sts)}

    # dominant kernel by HIP-event time: level-2 or level-k count or write
    def tot(key):
        return sum(s[key] for s in stats)
    kern = {
        "k_count2_batch": (tot("l2_count_kernel_ms"), tot("l2_count_bytes"), len(stats)),
        "k_countk_batch": (tot("count_kernel_ms") - tot("l2_count_kernel_ms"), tot("count_bytes") - tot("l2_count_bytes"),
                           tot("count_launches") - sum(1 for s in stats if s["l2_count_kernel_ms"] > 0)),
        "k_write2_batch": (tot("l2_write_kernel_ms"), tot("l2_write_bytes"),
                           sum(1 for s in stats if s["l2_write_kernel_ms"] > 0)),
        "k_writek_batch": (tot("write_kernel_ms") - tot("l2_write_kernel_ms"), tot("write_bytes") - tot("l2_write_bytes"),
                           tot("write_launches") - sum(1 for s in stats if s["l2_write_kernel_ms"] > 0)),
    }
    name = max(kern, key=lambda kn: kern[kn][0])
    ms, by, n = kern[name]
    achieved = (by / n) / ((ms / n) / 1000.0) / 1e9 if n and ms else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{name}_{args.workload}.json")
    if os.path.exists(tf) and args.N is None:
        with open(tf) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": name,
                       "launches": n, "avg_launch_ms": ms / n if n else None,
                       "algorithmic_bytes_per_launch": by / n if n else None}
    cnt_ms, wr_ms = tot("count_kernel_ms"), tot("write_kernel_ms")
    out["kernels_ms_per_step"] = {"count": cnt_ms / args.steps, "write": wr_ms / args.steps,
                                  "filter": sum(s["filter_kernel_ms"] for s in stats) / args.steps,
                                  "pi": sum(s["pi_kernel_ms"] for s in stats) / args.steps}

    if not args.no_plain:
        identical = True
        others = [colo.COLO_SCHED_PLAIN] + ([colo.COLO_SCHED_PERSISTENT] if main_sched != colo.COLO_SCHED_PERSISTENT
                                            else [])
        for sc in others:
            step(sc)  # warm
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = 1 if sc == colo.COLO_SCHED_PLAIN else args.steps
            for _ in range(reps):
                rep_o, _ = step(sc)
            torch.cuda.synchronize()
            out["mining_s"][sched_name[sc]] = (time.perf_counter() - t0) / reps
            identical &= rep_o == rep
            log(f"{sched_name[sc]}: {out['mining_s'][sched_name[sc]]:.3f} s per step")
        main_s = ms_per_step / 1000.0
        out["mining_s"][f"{sched_name[main_sched]}_speedup_vs_plain"] = out["mining_s"]["plain"] / main_s
        if "persistent" in out["mining_s"] and main_sched != colo.COLO_SCHED_PERSISTENT:
            out["mining_s"]["learned_speedup_vs_persistent"] = out["mining_s"]["persistent"] / main_s
        out["reports_identical"] = identical
        if not identical:
            log("ERROR: schedules produced different reports")
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(x, y, fc, bounds, p, args.cpu_budget, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
