#!/usr/bin/env python3
"""Trains the cost model of the learned schedule (c4 workload).

Draws randomised training datasets from the c2 generator (datagen.
training_config: features 10-40, 100k-2M instances, hotspot fraction 0-0.6,
sigma 0.1-2% of the side, r for 8-64 mean neighbours, theta 0.05-0.5; half
of them drawn wide: up to 60 features, the c3/c5 regimes, sizes up to 6; a
tenth with 2M-20M instances), generated by CPU worker processes while the
GPU mines the previous ones with the persistent schedule; the miner logs
(16 features; log1p rows, log1p refinement tests) for every counted
candidate.  Trains the
16-64-64-64-2 FP32 MLP with the HIP kernels of colo_mlp.hip, saves it, and
evaluates it on two held-out datasets: the 20M-instance, 40-feature c4
dataset (prediction errors, overflow fallbacks, end-to-end mining time of
the learned, persistent and plain schedules with identical reports) and a
2M-instance dataset inside the training range (the same, times as the best
of three runs).

    python gpu-colocationmininig_amd/train_cost_model.py --out models/cost_mlp.bin \\
        --datasets 200 --min-n 100000 --max-n 2000000 --eval-n 20000000
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import datagen  # noqa: E402


def make_dataset(job):
    """CPU worker: (seed, overrides) -> float32 arrays of one dataset."""
    seed, over = job
    over = dict(over)
    n = over.pop("N")
    labels, x, y, S, p = datagen.generate("c2", N=n, seed=seed, **over)
    fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
    return x.astype(np.float32), y.astype(np.float32), fc, S, p


def c4_eval_config(seed, n):
    cfg = datagen.training_config(np.random.default_rng(seed), n_range=(n, n))
    cfg["F"] = 40
    return cfg


def inrange_eval_config(seed):
    """A held-out dataset inside the training range (2M instances, wide)."""
    return datagen.training_config(np.random.default_rng(seed), n_range=(2_000_000, 2_000_000), f_range=(30, 50),
                                   wide=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "models", "cost_mlp.bin"))
    ap.add_argument("--datasets", type=int, default=200)
    ap.add_argument("--min-n", type=int, default=100_000)
    ap.add_argument("--max-n", type=int, default=2_000_000)
    ap.add_argument("--max-f", type=int, default=40, help="features per training dataset: 10 .. max-f")
    ap.add_argument("--wide-frac", type=float, default=0.5,
                    help="share of datasets drawn wide (10 .. wide-max-f features, sizes up to 6)")
    ap.add_argument("--wide-max-f", type=int, default=60)
    ap.add_argument("--big-frac", type=float, default=0.1,
                    help="share of datasets with max-n .. big-max-n instances")
    ap.add_argument("--big-max-n", type=int, default=20_000_000)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--seed", type=int, default=4000)
    ap.add_argument("--eval-n", type=int, default=20_000_000)
    ap.add_argument("--workers", type=int, default=12)
    ap.add_argument("--eval-only", action="store_true",
                    help="load --out and redo only the held-out evaluations (updates its .json)")
    a = ap.parse_args()
    if a.eval_only:
        return evaluate_only(a)
    rng = np.random.default_rng(a.seed)
    jobs = []
    for i in range(a.datasets):
        nr = (a.max_n, a.big_max_n) if rng.random() < a.big_frac else (a.min_n, a.max_n)
        if rng.random() < a.wide_frac:
            cfg = datagen.training_config(rng, n_range=nr, f_range=(10, a.wide_max_f), wide=True)
        else:
            cfg = datagen.training_config(rng, n_range=nr, f_range=(10, a.max_f))
        jobs.append((a.seed + i, cfg))
    # generation runs in spawned CPU workers (no GPU use there), started
    # before this process touches the GPU
    pool = mp.get_context("spawn").Pool(a.workers)
    it = pool.imap(make_dataset, jobs)
    ev = pool.apply_async(make_dataset, ((a.seed + 999, c4_eval_config(a.seed + 999, a.eval_n)),))
    ev2 = pool.apply_async(make_dataset, ((a.seed + 998, inrange_eval_config(a.seed + 998)),))

    import colo
    tmp = tempfile.mkdtemp(prefix="colo_train_")
    log = os.path.join(tmp, "train.bin")
    t0 = time.time()
    skipped = 0
    for i, (x, y, fc, S, p) in enumerate(it):
        ctx = colo.Context(x, y, fc, [0, 0, S, S], p["r"])
        try:
            rep, st = ctx.mine(p["theta"], max_k=p["max_k"], sched=colo.COLO_SCHED_PERSISTENT, hash_tables=False,
                               log_path=log)
        except colo.ColoError as e:
            skipped += 1
            print(f"[train] dataset {i + 1}: skipped ({e})", flush=True)
            ctx.close()
            continue
        ctx.close()
        print(f"[train] dataset {i + 1}/{a.datasets}: N={len(x)} F={p['F']} r={p['r']:g} "
              f"theta={p['theta']:.3f} max_k={p['max_k']} patterns={st['patterns']} ({time.time() - t0:.0f} s)",
              flush=True)
    X, T, K = colo.read_log(log, with_k=True)
    m = colo.CostModel(seed=a.seed)
    loss = m.train(X, T, steps=a.steps, batch=512, lr=2e-3, seed=a.seed)
    err = np.mean(np.abs(m.predict(X) - T), axis=0)
    print(f"[train] trained on {len(X)} records (sizes {np.bincount(K).tolist()}): loss {loss:.4f}, "
          f"mean |log err| rows {err[0]:.3f}, tests {err[1]:.3f} ({time.time() - t0:.0f} s)", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    m.save(a.out)

    summary = {"records": int(len(X)), "records_per_size": {int(k): int(v) for k, v in
                                                             enumerate(np.bincount(K)) if v},
               "datasets": a.datasets, "wide_frac": a.wide_frac, "skipped": skipped, "final_loss": loss,
               "train_mean_abs_log_err_rows": float(err[0]), "train_mean_abs_log_err_tests": float(err[1]),
               "model": os.path.basename(a.out), "outputs": colo.OUTPUT_NAMES}
    x, y, fc, S, p = ev.get()
    summary["c4_heldout"] = held_out(m, x, y, fc, S, p, reps=1)
    print(json.dumps(summary["c4_heldout"]), flush=True)
    x, y, fc, S, p = ev2.get()
    pool.close()
    pool.join()
    summary["inrange_heldout"] = held_out(m, x, y, fc, S, p, reps=3)
    print(json.dumps(summary["inrange_heldout"]), flush=True)
    summary["seconds"] = time.time() - t0
    with open(os.path.splitext(a.out)[0] + ".json", "w") as fh:
        json.dump(summary, fh, indent=1)


def held_out(m, x, y, fc, S, p, reps):
    """Held-out evaluation: both outputs' errors on the persistent run's
    log, and the learned / persistent / plain mining times (best of reps
    runs after a warm-up, default launch shapes, identical reports)."""
    import colo
    import torch
    tmp = tempfile.mkdtemp(prefix="colo_eval_")
    ctx = colo.Context(x, y, fc, [0, 0, S, S], p["r"])
    elog = os.path.join(tmp, "eval.bin")
    mk = p["max_k"]
    rep_p, _ = ctx.mine(p["theta"], max_k=mk, sched=colo.COLO_SCHED_PERSISTENT, hash_tables=False, log_path=elog)
    Xe, Te = colo.read_log(elog)
    err = np.mean(np.abs(m.predict(Xe) - Te), axis=0) if len(Xe) else [None, None]
    times, reps_out, lst = {}, {}, None
    for name, sched, model in (("learned", colo.COLO_SCHED_LEARNED, m), ("persistent", colo.COLO_SCHED_PERSISTENT, None),
                               ("plain", colo.COLO_SCHED_PLAIN, None)):
        ctx.mine(p["theta"], max_k=mk, sched=sched, model=model, hash_tables=False)  # warm
        best = None
        for _ in range(reps if name != "plain" else 1):
            torch.cuda.synchronize()
            ts = time.perf_counter()
            reps_out[name], st = ctx.mine(p["theta"], max_k=mk, sched=sched, model=model, hash_tables=False)
            torch.cuda.synchronize()
            dt = time.perf_counter() - ts
            best = dt if best is None else min(best, dt)
        times[name] = best
        if name == "learned":
            lst = st
    ctx.close()
    return {"records": int(len(Xe)), "mean_abs_log_err_rows": float(err[0]) if len(Xe) else None,
            "mean_abs_log_err_tests": float(err[1]) if len(Xe) else None,
            "instances": int(len(x)), "features": int(p["F"]), "theta": p["theta"], "r": p["r"], "max_k": mk,
            "learned_mean_abs_log_err_rows": lst["model_abs_err"] / max(lst["model_n"], 1),
            "overflow_fallbacks": lst["overflows"], "arenas": lst["arenas"], "mining_s": times,
            "learned_gain_over_persistent": times["persistent"] / times["learned"],
            "reports_identical": reps_out["learned"] == reps_out["persistent"] == reps_out["plain"] == rep_p}


def evaluate_only(a):
    import colo
    t0 = time.time()
    m = colo.CostModel.load(a.out)
    js = os.path.splitext(a.out)[0] + ".json"
    summary = {}
    if os.path.exists(js):
        with open(js) as fh:
            summary = json.load(fh)
    x, y, fc, S, p = make_dataset((a.seed + 999, c4_eval_config(a.seed + 999, a.eval_n)))
    summary["c4_heldout"] = held_out(m, x, y, fc, S, p, reps=1)
    print(json.dumps(summary["c4_heldout"]), flush=True)
    x, y, fc, S, p = make_dataset((a.seed + 998, inrange_eval_config(a.seed + 998)))
    summary["inrange_heldout"] = held_out(m, x, y, fc, S, p, reps=3)
    print(json.dumps(summary["inrange_heldout"]), flush=True)
    summary["seconds"] = time.time() - t0
    with open(js, "w") as fh:
        json.dump(summary, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
