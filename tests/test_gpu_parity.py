"""GPU parity: HIP kernels (through the C ABI) against the CPU oracle.

Bit-exact for ids, counts, offsets, tables, numerators; PI compared as
binary32 bit patterns.  Every case prints its name before touching the GPU.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _workload(datagen, name, N=None, seed=None, **over):
    labels, x, y, S, p = datagen.generate(name, N=N, seed=seed, **over)
    fcount = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
    bounds = np.array([0, 0, S, S], np.float32)
    return x.astype(np.float32), y.astype(np.float32), fcount, bounds, p


@pytest.fixture(scope="module")
def c1(datagen, colo_mod, oracle_mod):
    _say("fixture c1 N=20000")
    x, y, fc, b, p = _workload(datagen, "c1")
    ctx = colo_mod.Context(x, y, fc, b, p["r"])
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    orc.bin(p["r"])
    return ctx, orc, p


@pytest.fixture(scope="module")
def c2small(datagen, colo_mod, oracle_mod):
    _say("fixture c2 N=60000")
    x, y, fc, b, p = _workload(datagen, "c2", N=60000)
    ctx = colo_mod.Context(x, y, fc, b, p["r"])
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    orc.bin(p["r"])
    return ctx, orc, p


def _binning_expect(orc, fcount):
    cid, _ = orc.binning()
    n = len(cid)
    fid = np.repeat(np.arange(len(fcount)), fcount)
    order = np.lexsort((np.arange(n), cid, fid))  # (feature, cell, input index)
    return cid, order, fid


@pytest.mark.parametrize("fx", ["c1", "c2small"])
def test_binning(fx, request):
    ctx, orc, p = request.getfixturevalue(fx)
    _say(f"binning {fx}")
    g = ctx.grid()
    gx, gy, cells, r2 = orc.grid()
    assert (g["gx"], g["gy"], g["cells"]) == (gx, gy, cells)
    assert np.float32(g["r2"]).tobytes() == np.float32(r2).tobytes()
    cell, sidx, d = ctx.binning()
    cid, order, fid = _binning_expect(orc, ctx.fcount)
    np.testing.assert_array_equal(cell, cid)
    np.testing.assert_array_equal(sidx, order)
    # directory: first sorted position of every (feature, cell)
    key = fid.astype(np.int64) * (cells + 1) + cid
    expect = np.searchsorted(np.sort(key), np.arange(ctx.F * (cells + 1)), side="left")
    np.testing.assert_array_equal(d, expect.astype(np.uint32))


@pytest.mark.parametrize("case", ["one", "corners", "last_cell", "dense_cell"])
def test_directory_edge_cases(case, colo_mod, oracle_mod):
    """The directory on sparse layouts: a single instance, instances in the
    grid's corners (gaps of many 4096-entry tiles between run starts), every
    instance of a feature in the last cell, one cell holding every instance."""
    _say(f"directory {case}")
    S, r = 4000.0, 10.0
    if case == "one":
        x, y, fc = np.array([1234.5], np.float32), np.array([777.0], np.float32), np.array([1], np.uint32)
    elif case == "corners":
        x = np.array([0.5, S - 0.5, 0.5, S - 0.5, 2000.0], np.float32)
        y = np.array([0.5, 0.5, S - 0.5, S - 0.5, 2000.0], np.float32)
        fc = np.array([2, 1, 2], np.uint32)
    elif case == "last_cell":
        x = np.array([5.0, S - 0.1, S - 0.2, S - 0.3], np.float32)
        y = np.array([5.0, S - 0.1, S - 0.2, S - 0.3], np.float32)
        fc = np.array([1, 3], np.uint32)
    else:
        rng = np.random.default_rng(3)
        x = rng.uniform(1000.0, 1009.0, 600).astype(np.float32)
        y = rng.uniform(1000.0, 1009.0, 600).astype(np.float32)
        fc = np.array([300, 200, 100], np.uint32)
    b = np.array([0, 0, S, S], np.float32)
    ctx = colo_mod.Context(x, y, fc, b, r)
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    orc.bin(r)
    cells = ctx.grid()["cells"]
    cell, sidx, d = ctx.binning()
    cid, order, fid = _binning_expect(orc, fc)
    np.testing.assert_array_equal(cell, cid)
    np.testing.assert_array_equal(sidx, order)
    key = fid.astype(np.int64) * (cells + 1) + cid
    expect = np.searchsorted(np.sort(key), np.arange(len(fc) * (cells + 1)), side="left")
    np.testing.assert_array_equal(d, expect.astype(np.uint32))
    rep, _ = ctx.mine(0.1, max_k=3)
    assert rep == orc.mine(r, 0.1, max_k=3)
    ctx.close()


@pytest.mark.parametrize("fx", ["c1", "c2small"])
def test_filter_level2(fx, request):
    ctx, orc, p = request.getfixturevalue(fx)
    _say(f"filter level 2 {fx}")
    F = ctx.F
    cands = np.array([[i, j] for i in range(F) for j in range(i + 1, F)], np.uint32)
    gc, gp = ctx.filter(cands, 2, p["theta"])
    oc, op = orc.filter(cands, 2, p["theta"])
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gp, op)


def test_filter_level3(c2small):
    ctx, orc, p = c2small
    _say("filter level 3 c2small")
    F = ctx.F
    rng = np.random.default_rng(7)
    cands = np.sort(np.array([rng.choice(F, 3, replace=False) for _ in range(200)]), axis=1).astype(np.uint32)
    gc, gp = ctx.filter(cands, 3, p["theta"])
    oc, op = orc.filter(cands, 3, p["theta"])
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gp, op)


def _pairs(F, n, seed):
    rng = np.random.default_rng(seed)
    allp = [(i, j) for i in range(F) for j in range(i + 1, F)]
    idx = rng.choice(len(allp), size=min(n, len(allp)), replace=False)
    return [allp[i] for i in idx]


@pytest.mark.parametrize("fx", ["c1", "c2small"])
def test_join2(fx, request):
    ctx, orc, p = request.getfixturevalue(fx)
    for a, b in _pairs(ctx.F, 12, 3):
        _say(f"join2 {fx} ({a},{b})")
        rows, slots, bits = ctx.join2_count(a, b)
        orows, oslots, obm, oex = orc.join2_count(a, b)
        assert rows == orows
        Na, Nb = int(ctx.fcount[a]), int(ctx.fcount[b])
        np.testing.assert_array_equal(ctx._np_u32(slots, Na), oslots)
        part = ctx.participation(bits, [a, b])
        np.testing.assert_array_equal(part, oex[: Na + Nb].astype(np.uint8))
        nc, ne, pc, pe = ctx.pi(bits, [a, b])
        opc, onc = orc.pi_compat(obm, [a, b])
        ope, one = orc.pi_exact(oex, [a, b])
        np.testing.assert_array_equal(nc, onc)
        np.testing.assert_array_equal(ne, one)
        assert np.float32(pc).tobytes() == np.float32(opc).tobytes()
        assert np.float32(pe).tobytes() == np.float32(ope).tobytes()
        table, tot = ctx.join2_write(a, b, slots)
        assert tot == rows
        np.testing.assert_array_equal(ctx._np_u32(table, 2 * tot), orc.join2_write(a, b, oslots))


def test_joink(c2small, colo_mod):
    ctx, orc, p = c2small
    torch = ctx.torch
    F = ctx.F
    # grow level-3 and level-4 candidates from the densest pairs
    fc = ctx.fcount
    top = [int(v) for v in np.argsort(-fc)[:6]]
    top.sort()
    trip = [(top[0], top[1], top[2]), (top[1], top[3], top[4]), (top[0], top[2], top[5])]
    for cand in trip:
        a, b, cc = cand
        _say(f"joink k=3 {cand}")
        rows, slots, _ = ctx.join2_count(a, b)
        par_t, R = ctx.join2_write(a, b, slots)
        _, oslots, _, _ = orc.join2_count(a, b)
        opar = orc.join2_write(a, b, oslots)
        rows3, slots3, bits3 = ctx.joink_count(par_t, R, 3, list(cand))
        orows3, oslots3, obm3 = orc.joink_count(opar, 3, list(cand))
        assert rows3 == orows3
        np.testing.assert_array_equal(ctx._np_u32(slots3, R), oslots3)
        part = ctx.participation(bits3, list(cand))
        tot = int(sum(int(fc[f]) for f in cand))
        np.testing.assert_array_equal(part, obm3[:tot].astype(np.uint8))
        nc, ne, pc, pe = ctx.pi(bits3, list(cand))
        opc, onc = orc.pi_compat(obm3, list(cand))
        ope, one = orc.pi_exact(obm3, list(cand))
        np.testing.assert_array_equal(nc, onc)
        np.testing.assert_array_equal(ne, one)
        assert np.float32(pc).tobytes() == np.float32(opc).tobytes()
        t3, tot3 = ctx.joink_write(par_t, R, 3, list(cand), slots3)
        ot3 = orc.joink_write(opar, 3, list(cand), oslots3)
        np.testing.assert_array_equal(ctx._np_u32(t3, 3 * tot3), ot3)
        # the same pass fed by the anchors' neighbour lists: level-2 table of
        # (cand[0], cand[2]) and its row offsets
        _say(f"joink k=3 neighbour lists {cand}")
        _, s_ac, _ = ctx.join2_count(a, cc)
        _, _, t_ac, off_ac = ctx.join2_write(a, cc, s_ac, lists=True)
        rn, sn, bn = ctx.joink_count_nbr(par_t, R, 3, list(cand), t_ac, off_ac)
        assert rn == orows3
        np.testing.assert_array_equal(ctx._np_u32(sn, R), oslots3)
        ncn, nen, pcn, pen = ctx.pi(bn, list(cand))
        np.testing.assert_array_equal(ncn, onc)
        np.testing.assert_array_equal(nen, one)
        tn, totn = ctx.joink_write_nbr(par_t, R, 3, list(cand), sn, t_ac, off_ac)
        np.testing.assert_array_equal(ctx._np_u32(tn, 3 * totn), ot3)
        if tot3 == 0:
            continue
        # one more level on top
        for d in range(F):
            if d > int(cand[2]):
                c4 = list(cand) + [d]
                _say(f"joink k=4 {c4}")
                r4, s4, b4 = ctx.joink_count(t3, tot3, 4, c4)
                or4, os4, ob4 = orc.joink_count(ot3, 4, c4)
                assert r4 == or4
                np.testing.assert_array_equal(ctx._np_u32(s4, tot3), os4)
                nc, ne, pc, pe = ctx.pi(b4, c4)
                opc, onc = orc.pi_compat(ob4, c4)
                np.testing.assert_array_equal(nc, onc)
                t4, tot4 = ctx.joink_write(t3, tot3, 4, c4, s4)
                np.testing.assert_array_equal(ctx._np_u32(t4, 4 * tot4), orc.joink_write(ot3, 4, c4, os4))
                break


@pytest.mark.parametrize("sched", [0, 1])
@pytest.mark.parametrize("fx", ["c1", "c2small"])
def test_mine_report(fx, sched, request):
    ctx, orc, p = request.getfixturevalue(fx)
    _say(f"mine {fx} sched={sched}")
    rep, st = ctx.mine(p["theta"], max_k=p["max_k"], sched=sched)
    ref = orc.mine(p["r"], p["theta"], max_k=p["max_k"])
    assert rep == ref
    assert st["patterns"] > 0


@pytest.mark.parametrize("pi_mode", [0, 1])
def test_mine_modes(c2small, pi_mode):
    ctx, orc, p = c2small
    _say(f"mine c2small pi_mode={pi_mode} no-filter")
    rep, _ = ctx.mine(0.15, max_k=4, pi_mode=pi_mode, filter_on=False)
    ref = orc.mine(p["r"], 0.15, max_k=4, pi_mode=pi_mode, filter_on=False)
    assert rep == ref


def test_pattern_files(c1, tmp_path):
    ctx, orc, p = c1
    _say("pattern files c1")
    g = str(tmp_path) + "/gpu_"
    o = str(tmp_path) + "/orc_"
    ctx.mine(p["theta"], max_k=p["max_k"], out_prefix=g)
    orc.mine(p["r"], p["theta"], max_k=p["max_k"], out_prefix=o)
    for k in (2, 3, 4):
        fg, fo = g + f"ColocationDegree_{k}.txt", o + f"ColocationDegree_{k}.txt"
        assert os.path.exists(fg) == os.path.exists(fo)
        if os.path.exists(fo):
            assert open(fg).read() == open(fo).read()


@pytest.mark.parametrize("gx_case", ["gx1", "gx2", "gx3"])
def test_narrow_grids(gx_case, colo_mod, oracle_mod):
    """Grids one, two and three cells wide: the reference's neighbour-cell
    wrap rules (U:35-53) drop every cell when gx == 1 and the side column
    when gx == 2."""
    _say(f"narrow grid {gx_case}")
    rng = np.random.default_rng(11)
    r = 10.0
    W = {"gx1": 9.5, "gx2": 19.0, "gx3": 29.0}[gx_case]
    F, n = 4, 800
    x = rng.uniform(0, W, n).astype(np.float32)
    y = rng.uniform(0, 200, n).astype(np.float32)
    fc = np.full(F, n // F, np.uint32)
    b = np.array([0, 0, W, 200], np.float32)
    ctx = colo_mod.Context(x, y, fc, b, r)
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    rep, _ = ctx.mine(0.1, max_k=4)
    assert rep == orc.mine(r, 0.1, max_k=4)


def test_seattle_subset(colo_mod, oracle_mod):
    """Real data: the reference's Seattle2012 dataset
    (tests/golden/seattle2012.csv.gz), r = 10, theta = 0.3, sizes <= 3."""
    import gzip
    path = os.path.join(os.path.dirname(__file__), "golden", "seattle2012.csv.gz")
    if not os.path.exists(path):
        pytest.skip("fixture missing")
    _say("seattle r=10 theta=0.3 max_k=3")
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        csv = os.path.join(d, "s.csv")
        with gzip.open(path, "rb") as fi, open(csv, "wb") as fo:
            fo.write(fi.read())
        hd = colo_mod.HostData(csv)
        orc = oracle_mod.Oracle.from_csv(csv)
    ctx = colo_mod.Context.from_host(hd, 10.0)
    rep, _ = ctx.mine(0.3, max_k=3)
    assert rep == orc.mine(10.0, 0.3, max_k=3)


@pytest.mark.parametrize("sched", [0, 1, 2])
def test_threshold_sweep(c2small, colo_mod, sched):
    """One sweep over several thresholds equals one mining run per threshold
    (theta-independent results are reused across the sweep), under every
    schedule (the learned one with the shipped cost model)."""
    ctx, orc, p = c2small
    _say(f"threshold sweep sched={sched}")
    model = None
    if sched == colo_mod.COLO_SCHED_LEARNED:
        model = colo_mod.CostModel.load(os.path.join(os.path.dirname(colo_mod.__file__), "models", "cost_mlp.bin"))
    thetas = [0.25, 0.1, 0.15]
    out = ctx.mine_sweep(thetas, max_k=4, sched=sched, model=model)
    for th, (rep, st) in zip(thetas, out):
        single, _ = ctx.mine(th, max_k=4, sched=sched, model=model)
        assert rep == single, th
    assert out[0][0] == orc.mine(p["r"], 0.25, max_k=4)


@pytest.mark.parametrize("tune", [dict(count_var=1, write_var=1), dict(blocks_per_cu=4, item_rows=64),
                                  dict(count_var=1, write_var=0, blocks_per_cu=16, item_rows=4096),
                                  dict(count_var=2, write_var=2), dict(count_var=3, write_var=1),
                                  dict(write_var=2, write_rows_k=[0xFFFFFFFF] * 2 + [16384, 64, 4096] + [0xFFFFFFFF] * 11),
                                  dict(write_var=3), dict(count_var=11, write_var=3, item_rows=100),
                                  dict(write_var=3, write_rows_k=[0xFFFFFFFF] * 2 + [1000, 64, 70000 // 2] + [0xFFFFFFFF] * 11),
                                  dict(count_var=0, count_var_k=[0xFFFFFFFF] * 3 + [3, 1, 2] + [0xFFFFFFFF] * 10),
                                  dict(count_var=4, count_var_k=[0xFFFFFFFF] * 3 + [5, 4, 5] + [0xFFFFFFFF] * 10),
                                  dict(item_rows=1024, item_rows_k=[0xFFFFFFFF] * 2 + [64, 256, 16384] + [0xFFFFFFFF] * 11),
                                  dict(count_var=4, item_rows=256), dict(count_var=0, item_rows=64),
                                  dict(count_var=6), dict(count_var=7, item_rows=256),
                                  dict(count_var=8), dict(count_var=9, item_rows=64),
                                  dict(slots_budget_kb=64), dict(count_var=7, slots_budget_kb=16, item_rows=256),
                                  dict(count_var=10), dict(count_var=11, item_rows=256), dict(count_var=12, item_rows=64),
                                  dict(count_var=13, item_rows=16384)])
def test_launch_shapes_identical(c2small, tune):
    """Every launch shape of the batched passes gives the oracle's report;
    so does a slots budget small enough that the largest candidates are
    counted without per-row slots and recounted for their tables."""
    ctx, orc, p = c2small
    _say(f"launch shape {tune}")
    old = ctx.tune
    try:
        ctx.tune = tune
        for sched in (1, 2):
            rep, st = ctx.mine(0.15, max_k=4, sched=sched)
            assert rep == orc.mine(p["r"], 0.15, max_k=4), (tune, sched)
            if tune.get("slots_budget_kb"):
                assert st["slots_deferred"] > 0 and st["recounts"] > 0, st
    finally:
        ctx.tune = old


def test_autotune_and_cache(c2small, tmp_path):
    """The autotuner picks valid shapes, stores them, and takes them from the
    cache on the next call; results stay identical."""
    ctx, orc, p = c2small
    _say("autotune")
    old = ctx.tune
    cache = str(tmp_path / "tune.txt")
    try:
        t1 = ctx.autotune(0.15, max_k=4, cache=cache)
        assert os.path.getsize(cache) > 0
        assert t1["count_var_k"][3] <= 13 and all(v <= 13 or v == 0xFFFFFFFF for v in t1["count_var_k"]), t1
        assert t1["item_rows_k"][2] in (1024, 4096, 16384), t1
        ctx.tune = old
        t2 = ctx.autotune(0.15, max_k=4, cache=cache)
        assert t1 == t2 and ctx.tune == t1
        with open(cache) as fh:
            assert len(fh.read().splitlines()) == 1
        rep, _ = ctx.mine(0.15, max_k=4)
        assert rep == orc.mine(p["r"], 0.15, max_k=4)
    finally:
        ctx.tune = old


def test_binning_graph_replays(datagen, colo_mod, oracle_mod):
    """The binning captured as a HIP graph and replayed by colo_bin, with
    eager mining and synchronisations between the replays: every replay gives
    the eager binning, and mining after each replay gives the oracle's report;
    a new r makes a new capture."""
    _say("binning graph replays")
    x, y, fc, b, p = _workload(datagen, "c2", N=40000, seed=31)
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    orc.bin(p["r"])
    expect = orc.mine(p["r"], 0.2, max_k=3)
    eager = colo_mod.Context(x, y, fc, b, p["r"])
    eager.set_graphs(False)
    eager.rebin(p["r"])
    ref = eager.binning()
    ctx = colo_mod.Context(x, y, fc, b, p["r"])
    caps0, reps0 = ctx.graph_info()
    for i in range(3):
        ctx.rebin(p["r"])
        got = ctx.binning()
        for a, e in zip(got, ref):
            np.testing.assert_array_equal(a, e)
        assert ctx.mine(0.2, max_k=3)[0] == expect, i
    caps, reps = ctx.graph_info()
    assert caps0 == 1 and caps == 1 and reps == reps0 + 3
    ctx.rebin(p["r"] * 1.5)  # a new grid is a new capture
    assert ctx.graph_info()[0] == 2
    ctx.rebin(p["r"])
    assert ctx.graph_info()[0] == 3
    for a, e in zip(ctx.binning(), ref):
        np.testing.assert_array_equal(a, e)
    assert eager.graph_info()[1] == 1  # only the capture at creation was launched
    assert eager.mine(0.2, max_k=3)[0] == expect


def test_full_c2_all_schedules(datagen, colo_mod, oracle_mod):
    """The whole c2 workload (500k instances, sizes up to 5) under every
    schedule, with table checksums, against the oracle's report."""
    _say("full c2, all schedules")
    x, y, fc, b, p = _workload(datagen, "c2")
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    orc.bin(p["r"])
    expect = orc.mine(p["r"], p["theta"], max_k=p["max_k"])
    ctx = colo_mod.Context(x, y, fc, b, p["r"])
    model_path = os.path.join(os.path.dirname(colo_mod.__file__), "models", "cost_mlp.bin")
    model = colo_mod.CostModel.load(model_path)
    for sched, m in ((colo_mod.COLO_SCHED_PLAIN, None), (colo_mod.COLO_SCHED_PERSISTENT, None),
                     (colo_mod.COLO_SCHED_LEARNED, model)):
        rep, st = ctx.mine(p["theta"], max_k=p["max_k"], sched=sched, model=m)
        assert rep == expect, sched


@pytest.mark.parametrize("sched", ["plain", "persistent", "learned"])
def test_cli_drop_in(datagen, colo_mod, oracle_mod, tmp_path, sched):
    """The drop-in executable reads the reference's config.txt, parameter.txt
    and CSV and writes ColocationDegree_<k>.txt files equal to the oracle's,
    then its timing line (cpp:1144)."""
    import subprocess
    _say(f"cli drop-in {sched}")
    w = datagen.write_workload(str(tmp_path / "case"), "c1", N=6000)
    exe = os.path.join(os.path.dirname(colo_mod.__file__), "colo_miner")
    r = subprocess.run([exe, w["config"], "--sched", sched, "--max-k", "4"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    # timing line (cpp:1144): a newline, then dataset, executable, theta and
    # seven times in ms, " , "-separated, numbers as cout prints a float
    assert r.stdout.startswith("\n"), repr(r.stdout)
    fields = r.stdout[1:].split(" , ")
    assert len(fields) == 10, fields
    assert fields[0] == w["csv"] and fields[1] == "colo_miner"
    assert fields[2] == "%g" % np.float32(w["params"]["theta"]), fields[2]
    times = [float(v) for v in fields[3:]]
    assert all(t >= 0 for t in times) and times[1] > 0, times  # transfer measured
    assert times[-1] >= max(times[:-1]), times  # end to end spans the others
    orc = oracle_mod.Oracle.from_csv(w["csv"])
    o = str(tmp_path / "orc_")
    orc.mine(w["params"]["r"], w["params"]["theta"], max_k=4, out_prefix=o)
    loc = str(tmp_path / "case") + "/"
    for k in (2, 3, 4):
        fg, fo = loc + f"ColocationDegree_{k}.txt", o + f"ColocationDegree_{k}.txt"
        assert os.path.exists(fg) == os.path.exists(fo), k
        if os.path.exists(fo):
            assert open(fg).read() == open(fo).read(), k


def _run_cli(colo_mod, config, cwd, *extra):
    import subprocess
    exe = os.path.join(os.path.dirname(colo_mod.__file__), "colo_miner")
    return subprocess.run([exe, config, *extra], capture_output=True, text=True, timeout=600, cwd=cwd)


@pytest.mark.parametrize("sched", ["plain", "learned"])
def test_cli_hand_case(colo_mod, tmp_path, sched):
    """The drop-in executable writes the hand-worked pattern files (labels
    7, 3, 7, 12, 5; compat numerators with two instances per feature) byte
    for byte."""
    from test_oracle import HAND_CSV, HAND_FILES
    _say(f"cli hand case {sched}")
    (tmp_path / "hand.csv").write_text(HAND_CSV)
    (tmp_path / "parameter.txt").write_text("2\n0.5\n1\n")
    (tmp_path / "config.txt").write_text(f"{tmp_path}/\n{tmp_path}/parameter.txt\n{tmp_path}/hand.csv\n./\nlog.csv\n")
    r = _run_cli(colo_mod, str(tmp_path / "config.txt"), str(tmp_path), "--sched", sched)
    assert r.returncode == 0, r.stderr
    for k, text in HAND_FILES.items():
        assert (tmp_path / f"ColocationDegree_{k}.txt").read_text() == text, k
    assert not (tmp_path / "ColocationDegree_5.txt").exists()


def test_seattle_testcase_full_depth(colo_mod, tmp_path):
    """The reference's TestCase as it ships: config.txt with relative paths
    (TestCase/config.txt:1-5), Data/parameter.txt (r = 10, theta = 0.3) and
    Data/Seattle2012.csv, run from the TestCase directory with no size cap.
    The pattern files equal the oracle's (tests/golden/seattle_*), and the
    report with table checksums equals the oracle's complete report."""
    import gzip
    import shutil
    _say("seattle testcase, full depth, drop-in executable")
    g = os.path.join(os.path.dirname(__file__), "golden")
    (tmp_path / "Data").mkdir()
    with gzip.open(os.path.join(g, "seattle2012.csv.gz"), "rb") as fi, open(tmp_path / "Data" / "Seattle2012.csv", "wb") as fo:
        shutil.copyfileobj(fi, fo)
    (tmp_path / "Data" / "parameter.txt").write_text("10\n0.3\n1\n")
    (tmp_path / "config.txt").write_text("./\n./Data/parameter.txt\n./Data/Seattle2012.csv\n./ExperimentLog/\nExperimentLog1.csv\n")
    r = _run_cli(colo_mod, "config.txt", str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert r.stdout.split(" , ")[0] == "\n./Data/Seattle2012.csv"
    want = sorted(f for f in os.listdir(g) if f.startswith("seattle_ColocationDegree_"))
    assert want, "golden pattern files missing: run tests/golden/make_golden.py seattle"
    got = sorted(f for f in os.listdir(tmp_path) if f.startswith("ColocationDegree_"))
    assert got == [f[len("seattle_"):] for f in want]
    for f in want:
        assert (tmp_path / f[len("seattle_"):]).read_text() == open(os.path.join(g, f)).read(), f
    _say("seattle testcase, full depth, report")
    with gzip.open(os.path.join(g, "seattle_report.txt.gz"), "rt") as fh:
        fh.readline()
        golden = fh.read()
    hd = colo_mod.HostData(str(tmp_path / "Data" / "Seattle2012.csv"))
    ctx = colo_mod.Context.from_host(hd, 10.0)
    for sched in (colo_mod.COLO_SCHED_PLAIN, colo_mod.COLO_SCHED_PERSISTENT):
        rep, _ = ctx.mine(0.3, max_k=0, sched=sched, hash_tables=True)
        assert rep == golden, sched


@pytest.mark.parametrize("fx", ["c1", "c2small"])
def test_filter_threshold_on_a_bound(fx, request):
    """theta exactly on a candidate's bound passes it, one float32 step
    above fails it, as in the oracle (threshold_Kernel, K:545-573)."""
    ctx, orc, p = request.getfixturevalue(fx)
    _say(f"filter threshold on a bound {fx}")
    F = ctx.F
    cands = np.array([[i, j] for i in range(F) for j in range(i + 1, F)], np.uint32)
    oc, _ = orc.filter(cands, 2, 0.0)
    ub = [min(np.float32(np.float64(c) / np.float64(np.float32(ctx.fcount[f]))) for c, f in zip(row, cand))
          for row, cand in zip(oc, cands)]
    u = sorted(ub)[len(ub) // 2]
    for th in (u, np.nextafter(np.float32(u), np.float32(2))):
        gc, gp = ctx.filter(cands, 2, float(th))
        oc2, op2 = orc.filter(cands, 2, float(th))
        np.testing.assert_array_equal(gc, oc2)
        np.testing.assert_array_equal(gp, op2)


@pytest.mark.parametrize("F", [70, 130, 200])
def test_many_features(datagen, colo_mod, oracle_mod, F):
    """More than 64 features: the filter's feature masks span 2-4 64-bit
    words.  Filter counts and pass flags of pairs and triples (features on
    both sides of every 64-feature boundary among them), and the complete
    report of a mining run whose planted patterns use features above 64,
    equal the oracle's."""
    _say(f"many features F={F}")
    x, y, fc, b, p = _workload(datagen, "c1", N=40 * F, seed=F, F=F, planted=[3, 4], plant_rate=0.8,
                               plant_band=(F - 8, F), r=60.0, theta=0.3, max_k=4)
    ctx = colo_mod.Context(x, y, fc, b, p["r"])
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    orc.bin(p["r"])
    rng = np.random.default_rng(F)
    for k in (2, 3):
        cands = np.sort(np.array([rng.choice(F, k, replace=False) for _ in range(300)]), axis=1).astype(np.uint32)
        cands[:4, -1] = [63, 64, F - 1, 127 if F > 127 else F - 2]  # word boundaries
        cands = np.unique(np.sort(cands, axis=1), axis=0)
        cands = cands[np.all(np.diff(cands, axis=1) > 0, axis=1)]
        gc, gp = ctx.filter(cands, k, p["theta"])
        oc, op = orc.filter(cands, k, p["theta"])
        np.testing.assert_array_equal(gc, oc)
        np.testing.assert_array_equal(gp, op)
    rep, _ = ctx.mine(p["theta"], max_k=p["max_k"], sched=colo_mod.COLO_SCHED_PERSISTENT)
    assert rep == orc.mine(p["r"], p["theta"], max_k=p["max_k"])
    lv3 = [ln for ln in rep.split("\n") if ln.startswith("level 3 ")]
    assert lv3 and not lv3[0].endswith(" prev 0"), lv3  # a prevalent triple of features above 64


# 14 features of 20k instances around 30 shared hotspots: every pair is
# prevalent at theta 0.15 and most of the 364 triples are not
DENSE = dict(N=20000, seed=3, F=14, hot_frac=0.5, hotspots=30, sigma=(300, 300), theta=0.15, r=80.0)


@pytest.fixture(scope="module")
def dense(datagen, colo_mod, oracle_mod):
    _say("fixture dense c1 variant N=20000 F=14")
    x, y, fc, b, p = _workload(datagen, "c1", **DENSE)
    ctx = colo_mod.Context(x, y, fc, b, p["r"])
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    orc.bin(p["r"])
    return ctx, orc, p


@pytest.mark.parametrize("sched", ["plain", "persistent", "learned"])
@pytest.mark.parametrize("pi_mode", [0, 1])
def test_subset_bound(dense, colo_mod, sched, pi_mode):
    """Mining with the subset bound: bounds, skipped candidates, counts and
    tables equal the oracle's under every schedule and numerator mode, and
    the prevalent patterns and tables equal those of mining without it."""
    ctx, orc, p = dense
    _say(f"subset bound sched={sched} pi_mode={pi_mode}")
    sc = {"plain": colo_mod.COLO_SCHED_PLAIN, "persistent": colo_mod.COLO_SCHED_PERSISTENT,
          "learned": colo_mod.COLO_SCHED_LEARNED}[sched]
    model = None
    if sc == colo_mod.COLO_SCHED_LEARNED:
        model = colo_mod.CostModel.load(os.path.join(os.path.dirname(colo_mod.__file__), "models", "cost_mlp.bin"))
    rep, st = ctx.mine(p["theta"], max_k=4, pi_mode=pi_mode, sched=sc, model=model, subset_bound=True)
    ref = orc.mine(p["r"], p["theta"], max_k=4, pi_mode=pi_mode, subset_bound=True)
    assert rep == ref
    assert st["pruned"] > 200
    full, _ = ctx.mine(p["theta"], max_k=4, pi_mode=pi_mode, sched=sc, model=model)
    from report_check import parse_report
    c_on, t_on, l_on = parse_report(rep)
    c_off, t_off, l_off = parse_report(full)
    assert t_on == t_off and l_on == l_off
    assert {c for c, d in c_on.items() if d.get("v")} == {c for c, d in c_off.items() if d.get("v")}


@pytest.mark.parametrize("sched", ["plain", "persistent"])
@pytest.mark.parametrize("hashes", [True, False])
def test_subset_bound_sweep(dense, colo_mod, sched, hashes):
    """A threshold sweep with the subset bound equals one mining run per
    threshold (bounds and participation are kept across the sweep); without
    table checksums only the tables that surviving candidates extend are
    built."""
    ctx, orc, p = dense
    _say(f"subset bound sweep sched={sched} hashes={hashes}")
    sc = colo_mod.COLO_SCHED_PLAIN if sched == "plain" else colo_mod.COLO_SCHED_PERSISTENT
    thetas = [0.3, 0.1, 0.15]
    out = ctx.mine_sweep(thetas, max_k=4, sched=sc, subset_bound=True, hash_tables=hashes)
    for th, (rep, st) in zip(thetas, out):
        assert rep == orc.mine(p["r"], th, max_k=4, subset_bound=True, hash_tables=hashes), th


def test_subset_bound_c3_small(datagen, colo_mod, oracle_mod):
    """c3's generator at 200k instances, sizes up to 5: the learned schedule
    with the subset bound against the oracle with it."""
    _say("subset bound c3 N=200000")
    x, y, fc, b, p = _workload(datagen, "c3", N=200000, seed=7)
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    orc.bin(p["r"])
    ctx = colo_mod.Context(x, y, fc, b, p["r"])
    model = colo_mod.CostModel.load(os.path.join(os.path.dirname(colo_mod.__file__), "models", "cost_mlp.bin"))
    rep, st = ctx.mine(p["theta"], max_k=p["max_k"], sched=colo_mod.COLO_SCHED_LEARNED, model=model,
                       subset_bound=True)
    assert rep == orc.mine(p["r"], p["theta"], max_k=p["max_k"], subset_bound=True)
    assert st["pruned"] > 500


@pytest.mark.parametrize("sched", ["persistent", "learned"])
def test_level_graphs(datagen, colo_mod, oracle_mod, sched):
    """Each level's count, PI, scan and write passes run as captured HIP graphs:
    three rounds of mining with graphs, each followed by eager mining on the
    same context, all equal to the oracle's report (with the subset bound
    and table checksums); graphs off launches none."""
    _say(f"level graphs sched={sched}")
    x, y, fc, b, p = _workload(datagen, "c1", **DENSE)
    orc = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    orc.bin(p["r"])
    expect = orc.mine(p["r"], p["theta"], max_k=4, subset_bound=True)
    ctx = colo_mod.Context(x, y, fc, b, p["r"])
    sc = colo_mod.COLO_SCHED_PERSISTENT if sched == "persistent" else colo_mod.COLO_SCHED_LEARNED
    model = None
    if sc == colo_mod.COLO_SCHED_LEARNED:
        model = colo_mod.CostModel.load(os.path.join(os.path.dirname(colo_mod.__file__), "models", "cost_mlp.bin"))
    for i in range(3):
        ctx.set_graphs(True)
        rep, st = ctx.mine(p["theta"], max_k=4, sched=sc, model=model, subset_bound=True)
        assert rep == expect, i
        assert st["graphs"] >= 4 * st["levels"] - 2, st["graphs"]
        # the timing marks inside the graphs are recorded
        assert st["count_kernel_ms"] > 0 and st["write_kernel_ms"] > 0 and st["pi_kernel_ms"] > 0
        ctx.set_graphs(False)
        rep, st = ctx.mine(p["theta"], max_k=4, sched=sc, model=model, subset_bound=True)
        assert rep == expect, i
        assert st["graphs"] == 0


@pytest.mark.parametrize("sched", ["plain", "persistent", "learned"])
def test_tables_for_surviving_candidates(dense, colo_mod, sched):
    """Without table checksums, a level builds only the tables that the next
    level's candidates passing its filter and subset bound extend: the report
    (which lists the reference's tables and their rows) equals the oracle's,
    with and without the subset bound."""
    ctx, orc, p = dense
    _say(f"tables for surviving candidates sched={sched}")
    sc = {"plain": colo_mod.COLO_SCHED_PLAIN, "persistent": colo_mod.COLO_SCHED_PERSISTENT,
          "learned": colo_mod.COLO_SCHED_LEARNED}[sched]
    model = None
    if sc == colo_mod.COLO_SCHED_LEARNED:
        model = colo_mod.CostModel.load(os.path.join(os.path.dirname(colo_mod.__file__), "models", "cost_mlp.bin"))
    for sb in (True, False):
        rep, st = ctx.mine(p["theta"], max_k=4, sched=sc, model=model, subset_bound=sb, hash_tables=False)
        assert rep == orc.mine(p["r"], p["theta"], max_k=4, subset_bound=sb, hash_tables=False), sb
        if sb:
            assert st["tables_skipped"] > 0
