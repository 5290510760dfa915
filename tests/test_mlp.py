"""Cost model: the numpy oracle is pinned by finite differences and by
closed-form gradients that fix the parameter layout (CPU); the HIP forward /
backward / Adam kernels are checked against it (GPU); the two targets the
miner logs (table rows and refinement tests) are recomputed by the oracle
(GPU); the learned schedule is checked for identical mining results, its
overflow fallback is forced from the host, and its gain over the schedule
without the model is measured and asserted on a held-out dataset."""
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
    t0 = np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.25 * X[:, 3] ** 2 + 0.1 * X[:, 4:].sum(1)
    t1 = np.cos(X[:, 5]) - 0.3 * X[:, 6] * X[:, 0] + 0.2 * np.abs(X[:, 7])
    return X, np.stack([t0, t1], 1)


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
    """Closed forms fix the layout.  All parameters zero: the outputs are 0,
    the loss is sum(T^2)/n and only b4[o] gets 2 mean(-T[:, o]).  With b3[5]
    = 1 as well, hidden unit 5 of layer 3 is 1 for every sample, so
    W4[o][5] (flat index oW4 + o*64 + 5, row-major out x in) gets the same
    gradient as b4[o] and every other weight none."""
    import mlp_oracle as M
    off = {}
    o = 0
    for name, shp in M.SHAPES:
        off[name] = o
        o += int(np.prod(shp))
    assert o == M.NPARAMS == 64 * 16 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 2 * 64 + 2
    p = np.zeros(M.NPARAMS)
    X = np.ones((4, 16))
    T = np.array([[1.0, 5.0], [2.0, 6.0], [3.0, 7.0], [4.0, 8.0]])
    g, l = M.grad(p, X, T, np.zeros(16), np.ones(16))
    assert l == pytest.approx(np.sum(T ** 2) / 4)
    assert g[-2] == pytest.approx(2 * np.mean(-T[:, 0])) and g[-1] == pytest.approx(2 * np.mean(-T[:, 1]))
    assert np.all(g[:-2] == 0)
    p[off["b3"] + 5] = 1.0
    g, _ = M.grad(p, X, T, np.zeros(16), np.ones(16))
    for out in range(2):
        assert g[off["W4"] + out * 64 + 5] == pytest.approx(2 * np.mean(-T[:, out]))
    nz = set(np.nonzero(g)[0])
    assert nz == {off["W4"] + 5, off["W4"] + 64 + 5, off["b4"], off["b4"] + 1}, sorted(nz)


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
    assert T.shape == (len(X), 2) and np.all(T[:, 1] >= T[:, 0] - 1e-5)  # every row is a passed test
    m = colo_mod.CostModel(seed=1)
    m.train(X, T, steps=3000, batch=256, lr=2e-3)
    err = float(np.mean(np.abs(m.predict(X) - T)))
    return m, err, len(X)


def _ulps_ok(got, ref, ulps=2):
    got = np.asarray(got, np.float32)
    ref = np.asarray(ref, np.float32)
    return np.abs(got - ref) <= ulps * np.spacing(np.maximum(np.abs(ref), np.float32(1)))


@pytest.mark.gpu
def test_logged_targets_match_oracle(colo_mod, datagen, oracle_mod, tmp_path):
    """Both logged targets of counted candidates against the oracle, to 2
    binary32 ulps of log1p (the log's precision): log1p(rows) from the
    oracle's own count, and the refinement tests recomputed from the
    oracle's grid and tables -- at size 2, the sum over cells of N_first
    times N_last in the cell's 3x3 neighbourhood under the reference's
    neighbour-cell rule (orc_nbr_cell); at size k, the sum over the parent
    table's rows of the first instance's degree in the (first, last) pair
    table, both tables written by the oracle.  The pair-work feature (the
    3x3 join work of (first, last), computed once per unordered pair) is
    checked the same way."""
    from report_check import OracleChains, parse_report
    print("[case] logged cost-model targets vs oracle", flush=True)
    labels, x, y, S, p = datagen.generate("c2", N=60000, seed=995)
    fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    ctx = colo_mod.Context(x32, y32, fc, [0, 0, S, S], p["r"])
    log = str(tmp_path / "targets.bin")
    rep, _ = ctx.mine(p["theta"], max_k=4, sched=colo_mod.COLO_SCHED_PLAIN, log_path=log)
    X, T, K = colo_mod.read_log(log, with_k=True)
    cands, _, _ = parse_report(rep)
    counted = [c for c, d in cands.items() if d["p"]]  # report order = log order
    assert len(counted) == len(X) and [len(c) for c in counted] == K.tolist()
    assert max(K) >= 3
    orc = oracle_mod.Oracle.from_arrays(x32, y32, fc, np.array([0, 0, S, S], np.float32))
    orc.bin(p["r"])
    gx, gy, cells, _ = orc.grid()
    F = p["F"]
    cnt = orc.cell_counts()[0].reshape(cells, F).astype(np.int64)
    occ = np.nonzero(cnt.sum(1))[0]
    nb = np.zeros((cells, F), np.int64)
    for cur in occ:
        for i in (-1, 0, 1):
            for j in (-1, 0, 1):
                n = oracle_mod.nbr_cell(gx, cells, int(cur), i, j)
                if n >= 0:
                    nb[cur] += cnt[n]
    chains = OracleChains(orc)
    rng = np.random.default_rng(3)
    deep = [u for u, c in enumerate(counted) if len(c) > 2]
    pick = set(u for u, c in enumerate(counted) if len(c) == 2) | \
        set(rng.choice(deep, min(40, len(deep)), replace=False).tolist())
    n2 = nk = 0
    for u in sorted(pick):
        c = counted[u]
        k = len(c)
        if k == 2:
            rows = orc.join2_count(*c)[0]
            tests = int((cnt[:, c[0]] * nb[:, c[1]]).sum())
            n2 += 1
        else:
            parent = chains.table(c[:-1]).reshape(-1, k - 1)
            rows = orc.joink_count(parent.reshape(-1), k, list(c))[0]
            pair = chains.table((c[0], c[-1])).reshape(-1, 2)
            deg = np.bincount(pair[:, 0], minlength=len(x32))
            tests = int(deg[parent[:, 0]].sum())
            nk += 1
        assert rows == cands[c]["rows"], c
        assert tests >= rows, (c, tests, rows)
        ref = np.float32([np.log1p(rows), np.log1p(tests)])
        assert _ulps_ok(T[u], ref).all(), (c, T[u], ref, rows, tests)
        # feature 11: the 3x3 join work of (first, last), the level-2 tests
        work = int((cnt[:, c[0]] * nb[:, c[-1]]).sum())
        assert _ulps_ok(X[u][11], np.float32(np.log1p(work))), (c, X[u][11], work)
    print(f"[targets] {n2} pairs and {nk} larger candidates match the oracle", flush=True)
    assert n2 > 10 and nk > 5


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
def test_learned_half_prediction_releases_reservation(colo_mod, datagen):
    """The committed model with its rows output lowered by log 2, 4, 8, ...
    (until one table arena is under-sized) predicts a fraction of the
    tables: a reservation that is large but short.  The exact allocation
    must replace the reservation, not add to it: results equal the
    persistent schedule's, and the pool's peak stays within a quarter of
    the persistent schedule's (holding both would add the reservation to
    the level's tables)."""
    print("[case] learned schedule with under-sized predictions", flush=True)
    ctx, p = _ctx(colo_mod, datagen, "c2", 120000, 997)
    base = colo_mod.CostModel.load(os.path.join(os.path.dirname(colo_mod.__file__), "models", "cost_mlp.bin"))
    rep_p, st_p = ctx.mine(p["theta"], max_k=4, sched=colo_mod.COLO_SCHED_PERSISTENT)
    for shift in (2.0, 4.0, 8.0, 16.0, 64.0):
        m = colo_mod.CostModel.load(os.path.join(os.path.dirname(colo_mod.__file__), "models", "cost_mlp.bin"))
        par = base.params.copy()
        par[-2] -= np.float32(np.log(shift))  # b4[0]: the log1p(rows) output
        m.params = par
        rep_l, st_l = ctx.mine(p["theta"], max_k=4, sched=colo_mod.COLO_SCHED_LEARNED, model=m)
        assert rep_l == rep_p, shift
        print(f"[short] rows / {shift:g}: overflows {st_l['overflows']} of {st_l['arenas']} arenas, reserved "
              f"{st_l['arena_bytes'] / 1e6:.1f} MB; peak learned {st_l['pool_peak_bytes'] / 1e6:.1f} MB, persistent "
              f"{st_p['pool_peak_bytes'] / 1e6:.1f} MB", flush=True)
        if st_l["overflows"]:
            break
    assert st_l["overflows"] >= 1
    assert st_l["pool_peak_bytes"] <= 1.25 * st_p["pool_peak_bytes"]


# Largest time of the learned schedule, as a fraction of the persistent
# schedule's, that the gain test accepts on its held-out dataset.
GAIN_BOUND = 0.95


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_learned_gain_measured(colo_mod):
    """Ablation with the committed model on a held-out dataset inside its
    training range (2M instances, seed 4998, drawn like the wide training
    sets; no training seed is used): the persistent pipeline with and
    without the model, each at the default launch shapes, best of five runs
    after a warm-up.  Reports identical, and the model must take at most
    GAIN_BOUND of the time without it."""
    import torch
    import train_cost_model as tcm
    print("[case] learned vs persistent ablation (held-out, in range)", flush=True)
    m = colo_mod.CostModel.load(os.path.join(os.path.dirname(colo_mod.__file__), "models", "cost_mlp.bin"))
    x, y, fc, S, p = tcm.make_dataset((4998, tcm.inrange_eval_config(4998)))
    ctx = colo_mod.Context(x, y, fc, [0, 0, S, S], p["r"])
    res = {}
    for name, sched, model in (("persistent", colo_mod.COLO_SCHED_PERSISTENT, None),
                               ("learned", colo_mod.COLO_SCHED_LEARNED, m)):
        ctx.mine(p["theta"], max_k=p["max_k"], sched=sched, model=model, hash_tables=False)  # warm
        best = None
        for _ in range(5):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rep, st = ctx.mine(p["theta"], max_k=p["max_k"], sched=sched, model=model, hash_tables=False)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        res[name] = (best, rep, st)
    ctx.close()
    assert res["learned"][1] == res["persistent"][1]
    st = res["learned"][2]
    tl, tp = res["learned"][0], res["persistent"][0]
    print(f"[ablation] N={len(x)} F={p['F']} max_k={p['max_k']}: persistent {tp * 1e3:.2f} ms, "
          f"learned {tl * 1e3:.2f} ms (gain {tp / tl:.3f}x), model |log err| "
          f"{st['model_abs_err'] / max(st['model_n'], 1):.3f}, overflows {st['overflows']}", flush=True)
    assert tl <= GAIN_BOUND * tp, (tl, tp)
