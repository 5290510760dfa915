"""Cost model: the numpy oracle is pinned by finite differences (CPU); the HIP
forward / backward / Adam kernels are checked against it (GPU); the learned
schedule is checked for identical mining results, its overflow fallback is
forced from the host, and its gain over the schedule without the model is
measured."""
import os
import sys
import time

import numpy as np
import pytest


def _rand_model(rng):
    import mlp_oracle as M
    p = rng.normal(0, 0.3, M.NPARAMS)
    return p


def _data(rng, n):
    X = rng.normal(0, 1, (n, 16))
    T = np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.25 * X[:, 3] ** 2 + 0.1 * X[:, 4:].sum(1)
    return X, T


def test_oracle_gradient_finite_differences():
    import mlp_oracle as M
    rng = np.random.default_rng(0)
    p = _rand_model(rng)
    X, T = _data(rng, 40)
    mean, std = X.mean(0), X.std(0) + 1e-6
    g, _ = M.grad(p, X, T, mean, std)
    idx = rng.choice(M.NPARAMS, 300, replace=False)
    eps = 1e-6
    for i in idx:
        d = np.zeros(M.NPARAMS)
        d[i] = eps
        fd = (M.loss(p + d, X, T, mean, std) - M.loss(p - d, X, T, mean, std)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_oracle_gradient_layout():
    """A plain least-squares case fixes the layout: with every hidden unit
    off except bias paths, only b4 gets the mean residual gradient."""
    import mlp_oracle as M
    p = np.zeros(M.NPARAMS)
    X = np.ones((4, 16))
    T = np.array([1.0, 2.0, 3.0, 4.0])
    g, l = M.grad(p, X, T, np.zeros(16), np.ones(16))
    assert l == pytest.approx(np.mean(T ** 2))
    assert g[-1] == pytest.approx(2 * np.mean(0 - T))
    assert np.all(g[:-1] == 0)


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle(colo_mod):
    import mlp_oracle as M
    print("[case] mlp forward/gradient vs oracle", flush=True)
    rng = np.random.default_rng(1)
    m = colo_mod.CostModel(seed=3)
    p = rng.normal(0, 0.2, M.NPARAMS).astype(np.float32)
    m.params = p
    X, T = _data(rng, 777)
    X, T = X.astype(np.float32), T.astype(np.float32)
    mean, std = X.mean(0).astype(np.float32), (X.std(0) + 1e-3).astype(np.float32)
    m.set_norm(mean, std)
    y = m.predict(X)
    yo, _ = M.forward(p.astype(np.float64), X, mean.astype(np.float64), std.astype(np.float64))
    np.testing.assert_allclose(y, yo, rtol=1e-4, atol=1e-4)
    g, loss = m.grad(X, T)
    go, lo = M.grad(p.astype(np.float64), X, T, mean.astype(np.float64), std.astype(np.float64))
    assert loss == pytest.approx(lo, rel=1e-4)
    scale = np.abs(go).max()
    np.testing.assert_allclose(g, go, rtol=2e-3, atol=1e-5 * scale)


@pytest.mark.gpu
def test_gpu_training_fits(colo_mod):
    print("[case] mlp training", flush=True)
    rng = np.random.default_rng(2)
    X, T = _data(rng, 4096)
    X, T = X.astype(np.float32), T.astype(np.float32)
    m = colo_mod.CostModel(seed=5)
    m.set_norm(X.mean(0), X.std(0))
    y0 = m.predict(X)
    l0 = float(np.mean((y0 - T) ** 2))
    m.train(X, T, steps=1500, batch=256, lr=2e-3, normalise=False)
    y1 = m.predict(X)
    l1 = float(np.mean((y1 - T) ** 2))
    assert l1 < 0.1 * l0, (l0, l1)


def _ctx(colo_mod, datagen, name, N, seed, **over):
    labels, x, y, S, p = datagen.generate(name, N=N, seed=seed, **over)
    fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
    return colo_mod.Context(x.astype(np.float32), y.astype(np.float32), fc, [0, 0, S, S], p["r"]), p


@pytest.fixture(scope="module")
def trained(colo_mod, datagen, tmp_path_factory):
    """A model trained on logs of small randomised c2-style datasets."""
    d = tmp_path_factory.mktemp("mlp")
    log = str(d / "train.bin")
    rng = np.random.default_rng(11)
    for i in range(12):
        cfg = datagen.training_config(rng, n_range=(30_000, 120_000))
        n = cfg.pop("N")
        print(f"[case] logging dataset {i}: N={n} F={cfg['F']}", flush=True)
        ctx, p = _ctx(colo_mod, datagen, "c2", n, 100 + i, **cfg)
        ctx.mine(p["theta"], max_k=4, sched=colo_mod.COLO_SCHED_PERSISTENT, hash_tables=False, log_path=log)
    X, T = colo_mod.read_log(log)
    assert len(X) > 500, len(X)
    m = colo_mod.CostModel(seed=1)
    m.train(X, T, steps=3000, batch=256, lr=2e-3)
    err = float(np.mean(np.abs(m.predict(X) - T)))
    return m, err, len(X)


@pytest.mark.gpu
def test_learned_schedule_identical_results(colo_mod, datagen, oracle_mod, trained):
    m, err, n = trained
    print(f"[case] learned schedule parity (train err {err:.3f} on {n} records)", flush=True)
    labels, x, y, S, p = datagen.generate("c2", N=60000, seed=999)
    fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    ctx = colo_mod.Context(x32, y32, fc, [0, 0, S, S], p["r"])
    rep, st = ctx.mine(p["theta"], max_k=4, sched=colo_mod.COLO_SCHED_LEARNED, model=m)
    orc = oracle_mod.Oracle.from_arrays(x32, y32, fc, np.array([0, 0, S, S], np.float32))
    assert rep == orc.mine(p["r"], p["theta"], max_k=4)
    assert st["model_n"] > 0
    assert st["model_abs_err"] / st["model_n"] < 2.0


@pytest.mark.gpu
def test_learned_overflow_fallback(colo_mod, datagen):
    """A model that predicts zero rows under-sizes every table arena: each
    level must take the exact allocation and give the same results."""
    print("[case] learned schedule overflow fallback", flush=True)
    ctx, p = _ctx(colo_mod, datagen, "c2", 60000, 998)
    tiny = colo_mod.CostModel(seed=0)
    tiny.params = np.zeros(colo_mod.MLP_NPARAMS, np.float32)
    rep_l, st = ctx.mine(p["theta"], max_k=4, sched=colo_mod.COLO_SCHED_LEARNED, model=tiny)
    rep_p, _ = ctx.mine(p["theta"], max_k=4, sched=colo_mod.COLO_SCHED_PERSISTENT)
    assert rep_l == rep_p
    assert st["overflows"] >= 1


@pytest.mark.gpu
def test_learned_gain_measured(colo_mod, datagen, trained):
    """Ablation on a held-out dataset: the same persistent pipeline with and
    without the model (identical reports; times printed)."""
    m, err, _ = trained
    print("[case] learned vs persistent ablation", flush=True)
    ctx, p = _ctx(colo_mod, datagen, "c2", 200000, 997)
    import torch
    res = {}
    for name, sched, model in (("persistent", colo_mod.COLO_SCHED_PERSISTENT, None),
                               ("learned", colo_mod.COLO_SCHED_LEARNED, m)):
        ctx.mine(p["theta"], max_k=5, sched=sched, model=model, hash_tables=False)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            rep, st = ctx.mine(p["theta"], max_k=5, sched=sched, model=model, hash_tables=False)
        torch.cuda.synchronize()
        res[name] = ((time.perf_counter() - t0) / 3, rep, st)
    assert res["learned"][1] == res["persistent"][1]
    st = res["learned"][2]
    print(f"[ablation] persistent {res['persistent'][0]*1e3:.2f} ms, learned {res['learned'][0]*1e3:.2f} ms, "
          f"model |log err| {st['model_abs_err'] / max(st['model_n'], 1):.3f}, overflows {st['overflows']}",
          flush=True)
