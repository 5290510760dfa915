"""GPU parity on the benchmark configurations and on deep pattern sizes.

- c3 (5M instances, sizes <= 5): every schedule's complete report, with
  instance-table checksums, equals the oracle's complete report, committed as
  tests/golden/c3_report.txt.gz by tests/golden/make_golden.py (the oracle
  needs minutes for it; the dataset fingerprint in its header guards against
  a generator change).
- c5 (50M instances, sizes <= 6, four thresholds): the three schedules give
  identical sweeps with table checksums; a seeded sample of candidates per
  level (the five largest pairs, random ones at every level, filtered-out
  ones) is recomputed by the oracle, parents built by its own write passes.
- c4 held-out evaluation (20M instances, 40 features): learned, persistent
  and plain schedules identical with checksums, a sampled oracle check, and
  the cost model's error and overflow count recorded.
- deep levels: a planted size-10 pattern mined to size 10 (the kernels'
  generic path above k = 8) under every count and write variant.
Every case prints its name before touching the GPU.
"""
import gzip
import os
import sys

import numpy as np
import pytest

from report_check import check_sample, choose_sample, dataset_fingerprint, parse_report

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _gen(datagen, name, N=None, seed=None, **over):
    labels, x, y, S, p = datagen.generate(name, N=N, seed=seed, **over)
    fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
    return x.astype(np.float32), y.astype(np.float32), fc, np.array([0, 0, S, S], np.float32), p


def _model(colo_mod):
    return colo_mod.CostModel.load(os.path.join(os.path.dirname(colo_mod.__file__), "models", "cost_mlp.bin"))


def _schedules(colo_mod, model):
    return (("learned", colo_mod.COLO_SCHED_LEARNED, model), ("persistent", colo_mod.COLO_SCHED_PERSISTENT, None),
            ("plain", colo_mod.COLO_SCHED_PLAIN, None))


@pytest.mark.timeout(600)
def test_c3_complete_against_oracle_golden(datagen, colo_mod):
    path = os.path.join(HERE, "golden", "c3_report.txt.gz")
    with gzip.open(path, "rt") as fh:
        head = fh.readline().split()
        golden = fh.read()
    _say("c3 generate")
    x, y, fc, b, p = _gen(datagen, "c3")
    assert head[head.index("sha256") + 1] == dataset_fingerprint(x, y, fc), \
        "the c3 generator changed: rerun tests/golden/make_golden.py c3"
    assert float(head[head.index("theta") + 1]) == p["theta"] and int(head[head.index("max_k") + 1]) == p["max_k"]
    ctx = colo_mod.Context(x, y, fc, b, p["r"])
    model = _model(colo_mod)
    for name, sched, m in _schedules(colo_mod, model):
        _say(f"c3 mine {name}")
        rep, st = ctx.mine(p["theta"], max_k=p["max_k"], sched=sched, model=m, hash_tables=True)
        assert rep == golden, name
    assert st["levels"] == 4 and st["patterns"] > 0


@pytest.mark.timeout(900)
def test_c5_sampled_oracle_parity(datagen, colo_mod, oracle_mod):
    _say("c5 generate")
    x, y, fc, b, p = _gen(datagen, "c5")
    ctx = colo_mod.Context(x, y, fc, b, p["r"])
    model = _model(colo_mod)
    reps = {}
    for name, sched, m in _schedules(colo_mod, model):
        _say(f"c5 sweep {name}")
        out = ctx.mine_sweep(p["thetas"], max_k=p["max_k"], sched=sched, model=m, hash_tables=True)
        reps[name] = [r for r, _ in out]
        if name == "learned":
            st = out[0][1]
            print(f"[c5] learned theta {p['thetas'][0]}: levels {st['levels']}, patterns {st['patterns']}, "
                  f"model |log err| {st['model_abs_err'] / max(st['model_n'], 1):.3f}, overflows {st['overflows']}, "
                  f"slots deferred {st['slots_deferred']}, recounts {st['recounts']}", flush=True)
    assert reps["learned"] == reps["persistent"] == reps["plain"]
    # the subset bound (learned schedule, table checksums): every level line,
    # table line and counted candidate's line as without it, the skipped
    # candidates not prevalent, and each bound at least the candidate's PI
    _say("c5 sweep learned with the subset bound")
    bound = [r for r, _ in ctx.mine_sweep(p["thetas"], max_k=p["max_k"], sched=colo_mod.COLO_SCHED_LEARNED,
                                          model=model, hash_tables=True, subset_bound=True)]
    n_skipped = 0
    for th, rb, rf in zip(p["thetas"], bound, reps["learned"]):
        cb, tb, lb = parse_report(rb)
        cf, tf, lf = parse_report(rf)
        assert tb == tf and lb == lf, th
        lines_f = {tuple(l.split(" f ")[0].split()[1:]): l for l in rf.splitlines() if l.startswith("c ")}
        for line in rb.splitlines():
            if not line.startswith("c "):
                continue
            key = tuple(line.split(" f ")[0].split()[1:])
            c = tuple(int(v) for v in key)
            t = line.split()
            has_q = "q" in t
            if has_q:
                q = np.uint32(int(t[t.index("q") + 1], 16)).view(np.float32)
                if "pic" in cf[c]:
                    assert np.uint32(int(cf[c]["pic"], 16)).view(np.float32) <= q, (th, c)
                del t[t.index("q"):t.index("q") + 2]
            if "rows" in cb[c] or not has_q:  # counted, or cut by the filter
                assert " ".join(t) == lines_f[key], (th, c)
            else:  # skipped by the bound
                assert cf[c]["v"] == 0, (th, c)
                n_skipped += 1
    print(f"[c5] subset bound: {n_skipped} candidate evaluations skipped over the sweep", flush=True)
    assert n_skipped > 1000
    ctx.close()
    # the stated depth: sizes 2..6 at the lowest threshold
    cands, tables, levels = parse_report(reps["learned"][0])
    assert len(levels) == 5, levels
    # theta-independent fields agree across the sweep's thresholds
    for th, rep in zip(p["thetas"][1:], reps["learned"][1:]):
        c2, _, _ = parse_report(rep)
        for c, d in c2.items():
            if c in cands and d["p"] and cands[c]["p"]:
                assert {k: v for k, v in d.items() if k != "v"} == {k: v for k, v in cands[c].items() if k != "v"}, c
                assert d["v"] == int(not (np.uint32(int(d["pic"], 16)).view(np.float32) < np.float32(th))), (th, c)
    _say("c5 oracle binning")
    oracle_mod.set_threads(16)
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    orc.bin(p["r"])
    for th, rep, seed in ((p["thetas"][0], reps["learned"][0], 5), (p["thetas"][-1], reps["learned"][-1], 6)):
        sample = choose_sample(rep, seed, per_level=None if seed == 5 else {2: 4, 3: 3, 4: 2, 5: 2, 6: 2})
        n_c, n_t = check_sample(oracle_mod, orc, rep, th, sample, say=_say)
        print(f"[c5] theta {th}: {n_c} candidates and {n_t} tables checked against the oracle", flush=True)
        assert n_c >= 10


@pytest.mark.timeout(600)
def test_c4_heldout_learned_vs_plain(datagen, colo_mod, oracle_mod):
    _say("c4 held-out generate (20M, 40 features)")
    cfg = datagen.training_config(np.random.default_rng(4999), n_range=(20_000_000, 20_000_000))
    cfg["F"] = 40
    n = cfg.pop("N")
    x, y, fc, b, p = _gen(datagen, "c2", N=n, seed=4999, **cfg)
    ctx = colo_mod.Context(x, y, fc, b, p["r"])
    model = _model(colo_mod)
    reps = {}
    for name, sched, m in _schedules(colo_mod, model):
        _say(f"c4 mine {name}")
        reps[name], st = ctx.mine(p["theta"], max_k=4, sched=sched, model=m, hash_tables=True)
        if name == "learned":
            err = st["model_abs_err"] / max(st["model_n"], 1)
            print(f"[c4] held-out: {st['model_n']:.0f} predictions, mean |log1p rows error| {err:.3f}, "
                  f"overflow fallbacks {st['overflows']} of {st['arenas']} arenas", flush=True)
            assert err < 1.0
    assert reps["learned"] == reps["persistent"] == reps["plain"]
    ctx.close()
    _say("c4 oracle sample")
    oracle_mod.set_threads(16)
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    orc.bin(p["r"])
    n_c, n_t = check_sample(oracle_mod, orc, reps["learned"], p["theta"],
                            choose_sample(reps["learned"], 4, per_level={2: 10, 3: 6, 4: 4}), say=_say)
    print(f"[c4] {n_c} candidates and {n_t} tables checked against the oracle", flush=True)
    assert n_c >= 5


def test_deep_levels_every_variant(datagen, colo_mod, oracle_mod):
    """A size-10 pattern planted at 90% in 12 uniform features: every subset
    is prevalent, so mining runs to size 10 (levels 9 and 10 use the generic
    kernels above k = 8).  Every count variant (with a write variant each)
    and both batched schedules give the oracle's report, with checksums."""
    _say("deep levels generate")
    x, y, fc, b, p = _gen(datagen, "c1", N=36000, seed=77, F=12, planted=[10], plant_rate=0.9,
                          plant_band=(0, 12), r=40.0, theta=0.3, max_k=0)
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    expect = orc.mine(p["r"], p["theta"], max_k=0)
    _, _, levels = parse_report(expect)
    assert len(levels) >= 9, levels
    ctx = colo_mod.Context(x, y, fc, b, p["r"])
    old = ctx.tune
    try:
        for var in range(14):
            ctx.tune = dict(count_var=var, write_var=var % 4, item_rows=256 if var % 2 else 1024)
            for sched in (colo_mod.COLO_SCHED_PERSISTENT, colo_mod.COLO_SCHED_PLAIN if var == 0 else None):
                if sched is None:
                    continue
                _say(f"deep levels count var {var} sched {sched}")
                rep, st = ctx.mine(p["theta"], max_k=0, sched=sched)
                assert rep == expect, (var, sched)
                assert st["levels"] == len(levels)
    finally:
        ctx.tune = old
