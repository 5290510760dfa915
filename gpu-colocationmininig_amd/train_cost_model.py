#!/usr/bin/env python3
"""Trains the cost model of the learned schedule (c4 workload).

Draws randomised training datasets from the c2 generator (datagen.
training_config: features 10-40, 100k-2M instances, hotspot fraction 0-0.6,
sigma 0.1-2% of the side, r for 8-64 mean neighbours, theta 0.05-0.5),
generated by CPU worker processes while the GPU mines the previous ones with
the persistent schedule; the miner logs (16 features, log1p rows) for every
counted candidate.  Trains the 16-64-64-64-1 FP32 MLP with the HIP kernels
of colo_mlp.hip, saves it, and evaluates it on a held-out 20M-instance,
40-feature dataset: prediction error, overflow fallbacks and end-to-end
mining time of the learned, persistent and plain schedules (identical
reports required).

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "models", "cost_mlp.bin"))
    ap.add_argument("--datasets", type=int, default=200)
    ap.add_argument("--min-n", type=int, default=100_000)
    ap.add_argument("--max-n", type=int, default=2_000_000)
    ap.add_argument("--max-f", type=int, default=40, help="features per training dataset: 10 .. max-f")
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--seed", type=int, default=4000)
    ap.add_argument("--eval-n", type=int, default=20_000_000)
    ap.add_argument("--workers", type=int, default=12)
    ap.add_argument("--eval-only", action="store_true",
                    help="load --out and redo only the held-out evaluation (updates its .json)")
    a = ap.parse_args()
    if a.eval_only:
        return evaluate_only(a)
    rng = np.random.default_rng(a.seed)
    jobs = []
    for i in range(a.datasets):
        jobs.append((a.seed + i, datagen.training_config(rng, n_range=(a.min_n, a.max_n), f_range=(10, a.max_f))))
    ecfg = datagen.training_config(np.random.default_rng(a.seed + 999), n_range=(a.eval_n, a.eval_n))
    ecfg["F"] = 40
    # generation runs in spawned CPU workers (no GPU use there), started
    # before this process touches the GPU
    pool = mp.get_context("spawn").Pool(a.workers)
    it = pool.imap(make_dataset, jobs)
    ev = pool.apply_async(make_dataset, ((a.seed + 999, ecfg),))

    import colo
    import torch
    tmp = tempfile.mkdtemp(prefix="colo_train_")
    log = os.path.join(tmp, "train.bin")
    t0 = time.time()
    skipped = 0
    for i, (x, y, fc, S, p) in enumerate(it):
        ctx = colo.Context(x, y, fc, [0, 0, S, S], p["r"])
        try:
            rep, st = ctx.mine(p["theta"], max_k=4, sched=colo.COLO_SCHED_PERSISTENT, hash_tables=False,
                               log_path=log)
        except colo.ColoError as e:
            skipped += 1
            print(f"[train] dataset {i + 1}: skipped ({e})", flush=True)
            ctx.close()
            continue
        ctx.close()
        print(f"[train] dataset {i + 1}/{a.datasets}: N={len(x)} F={p['F']} r={p['r']:g} "
              f"theta={p['theta']:.3f} patterns={st['patterns']} ({time.time() - t0:.0f} s)", flush=True)
    X, T = colo.read_log(log)
    m = colo.CostModel(seed=a.seed)
    loss = m.train(X, T, steps=a.steps, batch=512, lr=2e-3, seed=a.seed)
    train_err = float(np.mean(np.abs(m.predict(X) - T)))
    print(f"[train] trained on {len(X)} records: loss {loss:.4f}, mean |log err| {train_err:.3f} "
          f"({time.time() - t0:.0f} s)", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    m.save(a.out)

    # held-out evaluation
    x, y, fc, S, p = ev.get()
    pool.close()
    pool.join()
    summary = {"records": int(len(X)), "datasets": a.datasets, "skipped": skipped, "final_loss": loss,
               "train_mean_abs_log_err": train_err, "model": os.path.basename(a.out)}
    summary.update(held_out(a, m, x, y, fc, S, p, t0))
    print(json.dumps(summary), flush=True)
    with open(os.path.splitext(a.out)[0] + ".json", "w") as fh:
        json.dump(summary, fh, indent=1)


def held_out(a, m, x, y, fc, S, p, t0):
    """Held-out evaluation: model error on the persistent run's log, and the
    learned / persistent / plain mining times (identical reports)."""
    import colo
    import torch
    tmp = tempfile.mkdtemp(prefix="colo_eval_")
    ctx = colo.Context(x, y, fc, [0, 0, S, S], p["r"])
    elog = os.path.join(tmp, "eval.bin")
    rep_p, _ = ctx.mine(p["theta"], max_k=4, sched=colo.COLO_SCHED_PERSISTENT, hash_tables=False, log_path=elog)
    Xe, Te = colo.read_log(elog)
    times, reps, lst = {}, {}, None
    for name, sched, model in (("learned", colo.COLO_SCHED_LEARNED, m), ("persistent", colo.COLO_SCHED_PERSISTENT, None),
                               ("plain", colo.COLO_SCHED_PLAIN, None)):
        ctx.mine(p["theta"], max_k=4, sched=sched, model=model, hash_tables=False)  # warm
        torch.cuda.synchronize()
        ts = time.perf_counter()
        reps[name], st = ctx.mine(p["theta"], max_k=4, sched=sched, model=model, hash_tables=False)
        torch.cuda.synchronize()
        times[name] = time.perf_counter() - ts
        if name == "learned":
            lst = st
    return {"eval_records": int(len(Xe)),
            "eval_mean_abs_log_err": float(np.mean(np.abs(m.predict(Xe) - Te))) if len(Xe) else None,
            "eval_instances": int(len(x)), "eval_features": int(p["F"]), "eval_theta": p["theta"], "eval_r": p["r"],
            "eval_learned_mean_abs_log_err": lst["model_abs_err"] / max(lst["model_n"], 1),
            "eval_overflow_fallbacks": lst["overflows"], "eval_arenas": lst["arenas"], "eval_mining_s": times,
            "eval_reports_identical": reps["learned"] == reps["persistent"] == reps["plain"] == rep_p,
            "seconds": time.time() - t0}


def evaluate_only(a):
    import colo
    t0 = time.time()
    ecfg = datagen.training_config(np.random.default_rng(a.seed + 999), n_range=(a.eval_n, a.eval_n))
    ecfg["F"] = 40
    x, y, fc, S, p = make_dataset((a.seed + 999, ecfg))
    m = colo.CostModel.load(a.out)
    js = os.path.splitext(a.out)[0] + ".json"
    summary = {}
    if os.path.exists(js):
        with open(js) as fh:
            summary = json.load(fh)
    summary.update(held_out(a, m, x, y, fc, S, p, t0))
    print(json.dumps(summary), flush=True)
    with open(js, "w") as fh:
        json.dump(summary, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
