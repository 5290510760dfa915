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
size 3-6 (each a run of consecutive frequency ranks in 6..59, planted at 50%),
r for ~64 mean neighbours, patterns up to size 6, thresholds 0.1, 0.2, 0.3 and
0.5.  The data is generated in memory with a fixed seed (float64 coordinates
converted to float32, the reference's coordinate type) before the timed
region.  c1/c2/c3 are available with --workload.

Prints one JSON line.  value = neighbour pairs per second of the level-2
join (pairs counted by the level-2 count pass / wall time of the level-2
count + PI + scan + write); mining_s = wall time of a step.  Every schedule
mines with the subset bound (from size 3 on, candidates whose
participation-intersection bound is below theta are not counted; the
prevalent patterns and their tables are unchanged), and the timed schedule
also runs without it, with the two results compared.  The timed
schedule is the learned one (persistent launches whose work queue is ordered
and partitioned, and whose table arenas are pre-sized, by the trained cost
model in gpu-colocationmininig_amd/models/cost_mlp.bin) when that model is
present, else the persistent one.  The persistent schedule (autotuned for
itself) and the plain per-candidate schedule also run on the same data; one
more sweep of each schedule with table checksums must give identical reports.
The CPU baseline is the oracle (the C/OpenMP restatement of the reference) on
the host cores: level-2 joins of the same workload over a time budget, plus
complete mining of c2 (16 threads and 1 thread) and c1 beside the GPU's.
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
T_START = time.perf_counter()


def log(msg):
    print(f"[bench {time.perf_counter() - T_START:6.1f}s] {msg}", file=sys.stderr, flush=True)


def describe(p):
    th = ", ".join(f"{t:g}" for t in p["thetas"])
    s = f"{p['name']}: {p['N']} instances, {p['F']} features"
    if p.get("zipf"):
        s += f" (Zipf {p['zipf']:g})"
    if p.get("hotspots"):
        lo, hi = p["sigma"]
        s += (f", {p['hot_frac']:.0%} in {p['hotspots']} hotspots (sigma {lo:g}-{hi:g}"
              + (f", {p['hot_feats']} features per hotspot" if p.get("hot_feats") else "") + ")")
    sizes = sorted(set(len(x) for x in p["planted"]))
    s += f", {len(p['planted'])} planted patterns of size {sizes[0]}-{sizes[-1]} at {p['plant_rate']:.0%}"
    if p.get("plant_band"):
        s += f" (runs of consecutive frequency ranks in {p['plant_band'][0]}..{p['plant_band'][1] - 1})"
    s += f", r={p['r']:g}, theta in {{{th}}}, patterns up to size {p['max_k']}, seed {p['seed']}"
    return s


def gen(name, N=None):
    import datagen
    import numpy as np
    labels, x, y, S, p = datagen.generate(name, N=N)
    fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
    return x.astype(np.float32), y.astype(np.float32), fc, np.array([0, 0, S, S], np.float32), p


def cpu_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": aff,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_threads():
    # the GPU box gives one GPU's process 16 host threads, whatever nproc says
    return max(1, min(16, cpu_info()["affinity_cpus"]))


def cpu_level2_sample(x, y, fc, bounds, p, budget_s, threads):
    """The oracle's level-2 joins (count + write) of the same workload on
    host cores, over as many candidate pairs as fit the time budget."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.set_threads(threads)
    o = oracle.Oracle.from_arrays(x, y, fc, bounds)
    o.bin(p["r"])
    F = o.F
    pairs = [(a, b) for a in range(F) for b in range(a + 1, F)]
    pairs.sort(key=lambda ab: (ab[0] + ab[1]) % 7)  # spread over frequent and rare features
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
    return {"value": rows / t_tot if t_tot else 0.0, "unit": "neighbor pairs/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C/OpenMP restatement of the reference) level-2 joins, count + write, of {used} of "
                      f"{len(pairs)} feature pairs of {p['name']} ({rows} pairs in {t_tot:.1f} s on {threads} threads)",
            "sample_s": t_tot, "sample_pairs": used, "all_pairs": len(pairs)}


def essentials(report):
    """What the subset bound must leave unchanged: the level lines, the
    instance tables (rows and checksums) and the prevalent candidates' lines
    (bound field removed)."""
    keep = []
    for line in report.splitlines():
        if line.startswith("c ") and line.endswith(" v 1"):
            t = line.split()
            if "q" in t:
                i = t.index("q")
                del t[i:i + 2]
            keep.append(" ".join(t))
        elif not line.startswith("c "):
            keep.append(line)
    return keep


def cpu_vs_gpu_mining(colo, model, threads, subset_bound):
    """Complete mining of c2 (oracle on `threads` threads and on 1 thread) and
    c1 (1 thread), each beside the GPU learned schedule on the same data,
    reports (with table checksums) compared: in full without the subset
    bound, else their prevalent patterns and tables (the oracle runs the
    reference's candidate set either way)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import torch
    out = {}
    for name, tlist in (("c2", (threads, 1)), ("c1", (1,))):
        x, y, fc, b, p = gen(name)
        ctx = colo.Context(x, y, fc, b, p["r"])
        sched = colo.COLO_SCHED_LEARNED if model is not None else colo.COLO_SCHED_PERSISTENT
        kw = dict(max_k=p["max_k"], sched=sched, model=model, subset_bound=subset_bound)
        ctx.mine(p["theta"], hash_tables=False, **kw)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.rebin(p["r"])
        rep, _ = ctx.mine(p["theta"], hash_tables=False, **kw)
        torch.cuda.synchronize()
        g_s = time.perf_counter() - t0
        grep, _ = ctx.mine(p["theta"], hash_tables=True, **kw)
        entry = {"workload": describe(p), "gpu_s": g_s, "gpu_subset_bound": subset_bound}
        for t in tlist:
            oracle.set_threads(t)
            t0 = time.perf_counter()
            o = oracle.Oracle.from_arrays(x, y, fc, b)
            orep = o.mine(p["r"], p["theta"], max_k=p["max_k"], hash_tables=True)
            entry[f"cpu_s_{t}t"] = time.perf_counter() - t0
            same = essentials(orep) == essentials(grep) if subset_bound else orep == grep
            entry["identical"] = entry.get("identical", True) and same
            del o
        out[name] = entry
        log(f"cpu vs gpu {name}: " + ", ".join(f"{k} {v:.3f}" for k, v in entry.items() if k.endswith("_s") or "_s_" in k)
            + f", identical {entry['identical']}")
        ctx.close()
    return out


def cpu_full(colo, model, threads, name, subset_bound):
    """Complete mining of one workload by the oracle on `threads` host
    threads (the reference's candidate set) and by the GPU (learned
    schedule, best of three after a warm-up), reports with instance-table
    checksums compared (prevalent patterns and tables with the subset
    bound)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import torch
    x, y, fc, b, p = gen(name)
    th = min(p["thetas"]) if "thetas" in p else p["theta"]
    ctx = colo.Context(x, y, fc, b, p["r"])
    sched = colo.COLO_SCHED_LEARNED if model is not None else colo.COLO_SCHED_PERSISTENT
    grep, _ = ctx.mine(th, max_k=p["max_k"], sched=sched, model=model, hash_tables=True, subset_bound=subset_bound)
    best = None
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.rebin(p["r"])
        ctx.mine(th, max_k=p["max_k"], sched=sched, model=model, hash_tables=False, subset_bound=subset_bound)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    ctx.close()
    log(f"cpu full {name}: gpu {best:.3f} s; oracle on {threads} threads ...")
    oracle.set_threads(threads)
    os.environ["ORC_PROGRESS"] = "1"  # the oracle notes its progress on stderr
    t0 = time.perf_counter()
    o = oracle.Oracle.from_arrays(x, y, fc, b)
    orep = o.mine(p["r"], th, max_k=p["max_k"], hash_tables=True)
    cpu_s = time.perf_counter() - t0
    same = essentials(orep) == essentials(grep) if subset_bound else orep == grep
    out = {"workload": describe(p), "theta": th, "gpu_s": best, "gpu_subset_bound": subset_bound,
           f"cpu_s_{threads}t": cpu_s, "cores": threads, "identical": same, "kind": "port", **cpu_info()}
    log(f"cpu full {name}: oracle {cpu_s:.1f} s, identical {out['identical']}")
    return out


def e2e_child(args):
    """One-shot mining in a fresh process, from host arrays to pattern files:
    context creation (upload and binning), launch shapes from the
    autotuner's cache, one mining run at the lowest threshold."""
    import tempfile
    import colo
    import torch
    x, y, fc, bounds, p = gen(args.workload, N=args.N)
    th = min(p["thetas"]) if "thetas" in p else p["theta"]
    model = None if args.no_model or not os.path.exists(args.model) else colo.CostModel.load(args.model)
    sched = colo.COLO_SCHED_LEARNED if model is not None else colo.COLO_SCHED_PERSISTENT
    torch.cuda.init()
    with tempfile.TemporaryDirectory() as td:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx = colo.Context(x, y, fc, bounds, p["r"])
        sb = args.subset_bound == "on"
        ctx.autotune(th, max_k=p["max_k"], sched=sched, model=model, cache=args.tune_cache, subset_bound=sb)
        _, st = ctx.mine(th, max_k=p["max_k"], sched=sched, hash_tables=False, out_prefix=os.path.join(td, ""),
                         model=model, subset_bound=sb)
        torch.cuda.synchronize()
        e2e = time.perf_counter() - t0
        files = sorted(f for f in os.listdir(td) if f.startswith("ColocationDegree_"))
    print(json.dumps({"s": e2e, "theta": th, "subset_bound": sb, "mining_ms": st["total_ms"], "alloc_ms": st["alloc_ms"],
                      "alloc_gb": st["alloc_bytes"] / 1e9,
                      "level_ms": {str(k): round(v, 2) for k, v in enumerate(st["level_ms"]) if v > 0},
                      "pattern_files": files, "upload_and_bin_included": True, "csv_parse_included": False,
                      "launch_shapes": "autotuner cache (" + os.path.basename(args.tune_cache) + ")"}), flush=True)
    ctx.close()
    return 0


def mine_once(args):
    """One mining run of the workload at its lowest threshold (learned
    schedule when the model is present, subset bound as --subset-bound), with
    the launch shapes from the autotuner's cache (a cache miss tunes first)."""
    import colo
    import torch
    x, y, fc, bounds, p = gen(args.workload, N=args.N)
    th = min(p["thetas"]) if "thetas" in p else p["theta"]
    model = None if args.no_model or not os.path.exists(args.model) else colo.CostModel.load(args.model)
    sched = colo.COLO_SCHED_LEARNED if model is not None else colo.COLO_SCHED_PERSISTENT
    sb = args.subset_bound == "on"
    ctx = colo.Context(x, y, fc, bounds, p["r"])
    ctx.autotune(th, max_k=p["max_k"], sched=sched, model=model, cache=args.tune_cache, subset_bound=sb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, st = ctx.mine(th, max_k=p["max_k"], sched=sched, model=model, hash_tables=False, subset_bound=sb)
    torch.cuda.synchronize()
    print(json.dumps({"workload": describe(p), "theta": th, "ms": (time.perf_counter() - t0) * 1e3,
                      "level_count_kernel_ms": {str(k): round(v, 2) for k, v in enumerate(st["level_count_kernel_ms"])
                                                if v > 0}}), flush=True)
    ctx.close()
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5")
    ap.add_argument("--N", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the persistent and plain schedules")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--model", default=os.path.join(ROOT, "gpu-colocationmininig_amd", "models", "cost_mlp.bin"))
    ap.add_argument("--no-model", action="store_true")
    ap.add_argument("--subset-bound", choices=["on", "off"], default="on",
                    help="mine with the subset bound (default on; the timed schedule also runs without it)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--tune-cache", default=os.path.join(os.path.expanduser("~"), ".cache", "colo_tune.txt"))
    ap.add_argument("--e2e-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-full", default=None, metavar="WORKLOAD",
                    help="only time complete mining of WORKLOAD by the oracle on the host cores beside the GPU")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy 3 times (traffic-counter calibration)")
    ap.add_argument("--mine-once", action="store_true",
                    help="only mine the workload once at its lowest threshold with the timed schedule and the "
                         "autotuner's cached launch shapes (a short program for hardware-counter runs)")
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
    if args.e2e_child:
        return e2e_child(args)
    if args.mine_once:
        return mine_once(args)
    if args.cpu_full:
        import colo
        model = None if args.no_model or not os.path.exists(args.model) else colo.CostModel.load(args.model)
        print(json.dumps(cpu_full(colo, model, cpu_threads(), args.cpu_full, args.subset_bound == "on")), flush=True)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (multi-GPU sharding is a design "
                                   "note only)", "n_gpus": args.gpus}))
        return 2

    x, y, fc, bounds, p = gen(args.workload, N=args.N)
    log(f"generated {describe(p)}")

    import colo
    import torch

    ctx = colo.Context(x, y, fc, bounds, p["r"])
    ctx.set_verbose(1)
    torch.cuda.synchronize()
    log("dataset resident and binned")

    model = None
    if not args.no_model and os.path.exists(args.model):
        model = colo.CostModel.load(args.model)
    main_sched = colo.COLO_SCHED_LEARNED if model is not None else colo.COLO_SCHED_PERSISTENT
    sched_name = {colo.COLO_SCHED_LEARNED: "learned", colo.COLO_SCHED_PERSISTENT: "persistent",
                  colo.COLO_SCHED_PLAIN: "plain"}
    thetas = p["thetas"]
    sb = args.subset_bound == "on"

    def tune_for(sched, m, bound=sb):
        if args.no_tune:
            return ctx.tune
        t0 = time.perf_counter()
        os.makedirs(os.path.dirname(args.tune_cache), exist_ok=True)
        t = ctx.autotune(min(thetas), max_k=p["max_k"], sched=sched, model=m, cache=args.tune_cache,
                         subset_bound=bound)
        log(f"autotuned {sched_name[sched]} launch shapes{'' if bound else ' (no subset bound)'} in "
            f"{time.perf_counter() - t0:.1f} s: {t}")
        return t

    default_tune = ctx.tune
    tune = tune_for(main_sched, model)

    bin_ms = []

    def step(sched, hash_tables=False, bound=sb):
        t0 = time.perf_counter()
        ctx.rebin(p["r"])
        torch.cuda.synchronize()
        bin_ms.append((time.perf_counter() - t0) * 1000.0)
        out = ctx.mine_sweep(thetas, max_k=p["max_k"], sched=sched, hash_tables=hash_tables,
                             model=model if sched == colo.COLO_SCHED_LEARNED else None, subset_bound=bound)
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
                      "subset_bound": sb, "launch_shapes": tune}}
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
    first = stats[:len(thetas)]
    out["patterns"] = {f"{t:g}": s["patterns"] for t, s in zip(thetas, first)}
    out["levels"] = {f"{t:g}": s["levels"] for t, s in zip(thetas, first)}
    out["level_ms"] = {f"{t:g}": {str(k): round(v, 2) for k, v in enumerate(s["level_ms"]) if v > 0}
                       for t, s in zip(thetas, first)}
    out["subset_bound"] = {"on": sb, "pruned": {f"{t:g}": s["pruned"] for t, s in zip(thetas, first)},
                           "bound_ms_per_step": sum(s["bound_ms"] for s in stats) / args.steps}
    out["memory"] = {"pool_peak_gb": max(s["pool_peak_bytes"] for s in stats) / 1e9,
                     "alloc_ms_per_step": sum(s["alloc_ms"] for s in stats) / args.steps,
                     "slots_deferred": sum(s["slots_deferred"] for s in first),
                     "recounts": sum(s["recounts"] for s in first)}

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
            tj = json.load(fh)
        if tj.get("workload_key") == describe(p):
            traffic = tj.get("hbm_bytes_per_launch")
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": name,
                       "launches": n, "avg_launch_ms": ms / n if n else None,
                       "algorithmic_bytes_per_launch": by / n if n else None}
    out["kernels_ms_per_step"] = {kn: v[0] / args.steps for kn, v in kern.items()}
    out["kernels_ms_per_step"].update({"filter": tot("filter_kernel_ms") / args.steps,
                                       "pi": tot("pi_kernel_ms") / args.steps,
                                       "offsets_scan": tot("scan_kernel_ms") / args.steps})

    if not args.no_plain:
        main_s = ms_per_step / 1000.0
        if sb:
            # the timed schedule without the subset bound (every candidate
            # that passes the filter is counted, as in the reference), with
            # launch shapes tuned for that; same prevalent patterns and tables
            ntune = tune_for(main_sched, model, bound=False)
            step(main_sched, bound=False)  # warm
            reps = min(args.steps, 3)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                step(main_sched, bound=False)
            torch.cuda.synchronize()
            nb = (time.perf_counter() - t0) / reps
            out["mining_s"][f"{sched_name[main_sched]}_no_subset_bound"] = nb
            out["subset_bound"]["speedup"] = nb / main_s
            out["subset_bound"]["no_bound_launch_shapes"] = ntune
            log(f"{sched_name[main_sched]} without the subset bound: {nb:.3f} s per step")
            with_b, _ = step(main_sched, hash_tables=True)
            without_b, _ = step(main_sched, hash_tables=True, bound=False)
            same = all(essentials(a) == essentials(b) for a, b in zip(with_b, without_b))
            out["subset_bound"]["patterns_and_tables_identical"] = same
            log(f"with and without the subset bound, prevalent patterns and table checksums identical: {same}")
            ctx.tune = tune
        # persistent with the launch shapes tuned for itself
        if main_sched != colo.COLO_SCHED_PERSISTENT:
            ctx.tune = default_tune
            ptune = tune_for(colo.COLO_SCHED_PERSISTENT, None)
            step(colo.COLO_SCHED_PERSISTENT)  # warm
            reps = min(args.steps, 3)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                step(colo.COLO_SCHED_PERSISTENT)
            torch.cuda.synchronize()
            out["mining_s"]["persistent"] = (time.perf_counter() - t0) / reps
            out["mining_s"]["persistent_launch_shapes"] = ptune
            out["mining_s"]["learned_speedup_vs_persistent"] = out["mining_s"]["persistent"] / main_s
            log(f"persistent (own shapes): {out['mining_s']['persistent']:.3f} s per step")
            ctx.tune = tune
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step(colo.COLO_SCHED_PLAIN)
        torch.cuda.synchronize()
        out["mining_s"]["plain"] = time.perf_counter() - t0
        out["mining_s"][f"{sched_name[main_sched]}_speedup_vs_plain"] = out["mining_s"]["plain"] / main_s
        log(f"plain: {out['mining_s']['plain']:.3f} s per step")
        # one sweep per schedule with table checksums: reports must agree
        ref, _ = step(main_sched, hash_tables=True)
        identical = True
        for sc in (colo.COLO_SCHED_PERSISTENT, colo.COLO_SCHED_PLAIN):
            if sc == main_sched:
                continue
            r2, _ = step(sc, hash_tables=True)
            identical &= r2 == ref
        out["reports_identical"] = identical
        out["reports_checked"] = "learned, persistent and plain sweeps with instance-table checksums"
        log(f"schedules identical (with table checksums): {identical}")
        if not identical:
            log("ERROR: schedules produced different reports")
    if not args.no_plain:
        # level-2 join of every feature pair with the filter off (the
        # filter-independent neighbour-pair rate; sizes up to 3 so that the
        # level-2 tables the next level extends are written, as in a step)
        _, st0 = ctx.mine(min(thetas), max_k=3, filter_on=False, sched=main_sched, hash_tables=False,
                          model=model if main_sched == colo.COLO_SCHED_LEARNED else None, subset_bound=sb)
        out["level2_filter_off"] = {"pairs": st0["l2_pairs"], "join_ms": st0["l2_join_ms"],
                                    "pairs_per_s": st0["l2_pairs"] / (st0["l2_join_ms"] / 1000.0)
                                    if st0["l2_join_ms"] else None, "candidates": st0["candidates"]}
        log(f"level 2 with the filter off: {st0['l2_pairs']} pairs in {st0['l2_join_ms']:.1f} ms")
        ctx.close()
        # end to end from host-resident arrays, as a one-shot run: a new
        # process creates the context (upload and binning), takes the launch
        # shapes from the autotuner's cache and mines once at the lowest
        # threshold, writing the ColocationDegree_<k>.txt files
        import subprocess
        torch.cuda.empty_cache()
        cmd = [sys.executable, os.path.abspath(__file__), "--e2e-child", "--workload", args.workload,
               "--tune-cache", args.tune_cache, "--model", args.model, "--subset-bound", args.subset_bound] + \
              (["--no-model"] if model is None else []) + \
              (["--N", str(args.N)] if args.N else [])
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        if r.returncode == 0 and r.stdout.strip():
            out["end_to_end"] = json.loads(r.stdout.strip().splitlines()[-1])
            log(f"end to end at theta {out['end_to_end']['theta']:g}: {out['end_to_end']['s']:.3f} s "
                f"({len(out['end_to_end']['pattern_files'])} pattern files)")
        else:
            log("end-to-end run failed: " + r.stderr[-500:])
    if not args.no_cpu:
        threads = cpu_threads()
        log(f"cpu baseline on {threads} threads")
        info = cpu_info()
        cb = cpu_level2_sample(x, y, fc, bounds, p, args.cpu_budget, threads)
        cb.update(info)
        # the GPU's whole step against the CPU's level-2 joins of a fraction
        # of the pairs: a lower bound on the CPU's mining time of the workload
        cb["cpu_mining_s_lower_bound"] = cb["sample_s"]
        cb["gpu_step_s"] = ms_per_step / 1000.0
        out["cpu_baseline"] = cb
        out["cpu_vs_gpu_mining"] = cpu_vs_gpu_mining(colo, model, threads, sb)
    log("done")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
