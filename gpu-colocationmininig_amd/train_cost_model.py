#!/usr/bin/env python3
"""Trains the cost model of the learned schedule (c4 workload).

Draws randomised training datasets from the c2 generator (datagen.
training_config: features 10-40, hotspot fraction 0-0.6, sigma 0.1-2% of the
side, r for 8-64 mean neighbours, theta 0.05-0.3), mines each on the GPU with
the persistent schedule while the miner logs (16 features, log1p rows) for
every counted candidate, then trains the 16-64-64-64-1 FP32 MLP with the HIP
kernels of colo_mlp.hip and saves it.  Also evaluates on a held-out dataset.

    python gpu-colocationmininig_amd/train_cost_model.py --out models/cost_mlp.bin \\
        --datasets 40 --min-n 100000 --max-n 1000000
"""
import argparse
import json
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import colo  # noqa: E402
import datagen  # noqa: E402


def mine_logged(name, N, seed, over, log, max_k=4):
    labels, x, y, S, p = datagen.generate(name, N=N, seed=seed, **over)
    fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
    ctx = colo.Context(x.astype(np.float32), y.astype(np.float32), fc, [0, 0, S, S], p["r"])
    rep, st = ctx.mine(p["theta"], max_k=max_k, sched=colo.COLO_SCHED_PERSISTENT, hash_tables=False, log_path=log)
    ctx.close()
    return st, p


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "models", "cost_mlp.bin"))
    ap.add_argument("--datasets", type=int, default=40)
    ap.add_argument("--min-n", type=int, default=100_000)
    ap.add_argument("--max-n", type=int, default=1_000_000)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--seed", type=int, default=4000)
    ap.add_argument("--eval-n", type=int, default=2_000_000)
    a = ap.parse_args()
    rng = np.random.default_rng(a.seed)
    tmp = tempfile.mkdtemp(prefix="colo_train_")
    log = os.path.join(tmp, "train.bin")
    t0 = time.time()
    for i in range(a.datasets):
        cfg = datagen.training_config(rng, n_range=(a.min_n, a.max_n))
        n = cfg.pop("N")
        st, p = mine_logged("c2", n, a.seed + i, cfg, log)
        print(f"[train] dataset {i + 1}/{a.datasets}: N={n} F={p['F']} r={p['r']:g} theta={p['theta']:.3f} "
              f"patterns={st['patterns']} ({time.time() - t0:.0f} s)", flush=True)
    X, T = colo.read_log(log)
    m = colo.CostModel(seed=a.seed)
    loss = m.train(X, T, steps=a.steps, batch=512, lr=2e-3, seed=a.seed)
    train_err = float(np.mean(np.abs(m.predict(X) - T)))
    # held-out evaluation: a larger dataset with its own seed
    elog = os.path.join(tmp, "eval.bin")
    cfg = datagen.training_config(np.random.default_rng(a.seed + 999), n_range=(a.eval_n, a.eval_n))
    n = cfg.pop("N")
    cfg["F"] = 40
    mine_logged("c2", n, a.seed + 999, cfg, elog)
    Xe, Te = colo.read_log(elog)
    eval_err = float(np.mean(np.abs(m.predict(Xe) - Te))) if len(Xe) else None
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    m.save(a.out)
    summary = {"records": int(len(X)), "datasets": a.datasets, "final_loss": loss,
               "train_mean_abs_log_err": train_err, "eval_records": int(len(Xe)),
               "eval_mean_abs_log_err": eval_err, "eval_instances": n, "seconds": time.time() - t0,
               "model": a.out}
    print(json.dumps(summary), flush=True)
    with open(os.path.splitext(a.out)[0] + ".json", "w") as fh:
        json.dump(summary, fh, indent=1)


if __name__ == "__main__":
    main()
