"""Pins the CPU oracle against facts that do not come from the oracle itself.

- neighbour cells: the literal getNeighborCellID (Codes/Utility_Inline.h:35-53)
  against the clipping rule derived from it, on every cell of narrow grids;
- grid and binning: closed-form binary32 cell ids, numpy's stable argsort,
  bincount;
- joins: brute-force all pairs / all triples on small inputs, and points
  placed at r - ulp, r and r + ulp;
- PI: hand-computed bitmaps, including N = 1 and the degree-2 shared word;
- filter: literal quadrant sweep vs the fast form; soundness;
- candidate generation: hand-made lists that exercise the first-failure break;
- miner: filter on == filter off; exact numerators never grow with the size.
"""
import os
from fractions import Fraction

import numpy as np
import pytest


def nbr_rule(gx, gy, cur, i, j):
    """Neighbour cell of anchor `cur` at offset (i, j), or -1: rows clipped
    to the grid; gx >= 3 clips columns; gx == 2 keeps only the anchor's
    column; gx == 1 keeps nothing."""
    ay, ax = divmod(cur, gx)
    ry, rx = ay + j, ax + i
    if gx == 1 or not (0 <= ry < gy) or not (0 <= rx < gx):
        return -1
    if gx == 2 and rx != ax:
        return -1
    return ry * gx + rx


@pytest.mark.parametrize("gx", [1, 2, 3, 4, 7])
@pytest.mark.parametrize("gy", [1, 2, 3, 5])
def test_neighbour_cells(oracle_mod, gx, gy):
    cells = gx * gy
    for cur in range(cells):
        for j in (-1, 0, 1):
            for i in (-1, 0, 1):
                assert oracle_mod.nbr_cell(gx, cells, cur, i, j) == nbr_rule(gx, gy, cur, i, j), (gx, gy, cur, i, j)


def _dataset(rng, n_per, F, S, hot=0.0):
    pts = []
    for f in range(F):
        p = rng.uniform(0, S, size=(n_per, 2))
        if hot:
            k = int(hot * n_per)
            p[:k] = rng.normal(S / 2, S / 20, size=(k, 2))
        pts.append(np.clip(p, 0, S * 0.999))
    xy = np.concatenate(pts).astype(np.float32)
    return xy[:, 0].copy(), xy[:, 1].copy(), np.full(F, n_per, np.uint32), np.array([0, 0, S, S], np.float32)


def _cells32(x, y, b, r):
    r32 = np.float32(r)
    gx = int(np.ceil(np.float32(b[2] - b[0]) / r32))
    i = ((x - np.float32(b[0])) / r32).astype(np.uint32)
    j = ((y - np.float32(b[1])) / r32).astype(np.uint32)
    return (np.uint32(gx) * j + i).astype(np.uint32), gx


def test_grid_and_binning(oracle_mod):
    rng = np.random.default_rng(0)
    x, y, fc, b = _dataset(rng, 500, 4, 1000.0)
    o = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    r = 37.0
    o.bin(r)
    gx, gy, cells, r2 = o.grid()
    r32 = np.float32(r)
    assert gx == int(np.ceil(np.float32(b[2] - b[0]) / r32))
    assert gy == int(np.ceil(np.float32(np.float32(b[3]) + r32 - np.float32(b[1])) / r32))
    assert np.float32(r2) == r32 * r32
    cid, sidx = o.binning()
    exp, _ = _cells32(x, y, b, r)
    np.testing.assert_array_equal(cid, exp)
    np.testing.assert_array_equal(sidx, np.argsort(cid, kind="stable"))
    cnt, st = o.cell_counts()
    fid = np.repeat(np.arange(4), fc)
    dense = np.bincount(cid.astype(np.int64) * 4 + fid, minlength=cells * 4)
    np.testing.assert_array_equal(cnt, dense)
    np.testing.assert_array_equal(st, np.concatenate([[0], np.cumsum(dense)[:-1]]))


def _d2(x, y, p, q):
    dx = x[q] - x[p]
    dy = y[q] - y[p]
    return dx * dx + dy * dy  # binary32, two roundings each, no FMA in numpy


def _brute_pairs(x, y, cid, gx, gy, r2, ga, gb):
    """All (g, q) with q in the anchor's neighbour cells and d2 <= r2, anchors
    in order, neighbours by (cell, index)."""
    rows = []
    for g in ga:
        ms = [q for q in gb if _d2(x, y, g, q) <= r2 and _adjacent(cid[g], cid[q], gx, gy)]
        ms.sort(key=lambda q: (cid[q], q))
        rows.extend((g, q) for q in ms)
    return rows


def _adjacent(ca, cb, gx, gy):
    for j in (-1, 0, 1):
        for i in (-1, 0, 1):
            if nbr_rule(gx, gy, int(ca), i, j) == int(cb):
                return True
    return False


@pytest.mark.parametrize("seed,S,r", [(1, 300.0, 25.0), (2, 60.0, 25.0), (3, 200.0, 7.5)])
def test_join2_bruteforce(oracle_mod, seed, S, r):
    rng = np.random.default_rng(seed)
    x, y, fc, b = _dataset(rng, 150, 3, S, hot=0.3)
    o = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    o.bin(r)
    gx, gy, cells, r2 = o.grid()
    cid, _ = o.binning()
    fs = np.concatenate([[0], np.cumsum(fc)[:-1]]).astype(np.int64)
    for a, bb in [(0, 1), (0, 2), (1, 2)]:
        ga = range(int(fs[a]), int(fs[a] + fc[a]))
        gb = range(int(fs[bb]), int(fs[bb] + fc[bb]))
        exp = _brute_pairs(x, y, cid, gx, gy, np.float32(r2), ga, gb)
        rows, slots, bm, ex = o.join2_count(a, bb)
        assert rows == len(exp)
        t = o.join2_write(a, bb, slots)
        assert [tuple(v) for v in t.reshape(-1, 2).tolist()] == exp
        # participation (exact layout: a at t, b at N_a + rank)
        Na = int(fc[a])
        pa = np.zeros(Na, np.uint32)
        pb = np.zeros(int(fc[bb]), np.uint32)
        for g, q in exp:
            pa[g - fs[a]] = 1
            pb[q - fs[bb]] = 1
        np.testing.assert_array_equal(ex[: Na + len(pb)], np.concatenate([pa, pb]))


def test_join3_bruteforce(oracle_mod):
    rng = np.random.default_rng(5)
    x, y, fc, b = _dataset(rng, 60, 3, 120.0, hot=0.5)
    r = 20.0
    o = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    o.bin(r)
    gx, gy, cells, r2 = o.grid()
    r2 = np.float32(r2)
    cid, _ = o.binning()
    fs = np.concatenate([[0], np.cumsum(fc)[:-1]]).astype(np.int64)
    _, s2, _, _ = o.join2_count(0, 1)
    par = o.join2_write(0, 1, s2)
    exp = []
    for g, h in par.reshape(-1, 2).tolist():
        ms = [q for q in range(int(fs[2]), int(fs[2] + fc[2]))
              if _adjacent(cid[g], cid[q], gx, gy) and _d2(x, y, g, q) <= r2 and _d2(x, y, h, q) <= r2]
        ms.sort(key=lambda q: (cid[q], q))
        exp.extend((g, h, q) for q in ms)
    rows, slots, bm = o.joink_count(par, 3, [0, 1, 2])
    assert rows == len(exp)
    t = o.joink_write(par, 3, [0, 1, 2], slots)
    assert [tuple(v) for v in t.reshape(-1, 3).tolist()] == exp
    # every row is a pairwise clique
    for g, h, q in exp:
        assert _d2(x, y, g, h) <= r2 and _d2(x, y, g, q) <= r2 and _d2(x, y, h, q) <= r2


def test_distance_boundary(oracle_mod):
    """Pairs at r - ulp, r and r + ulp (binary32) from an anchor at the
    origin, where every offset is exact."""
    r = np.float32(10.0)
    up = np.nextafter(r, np.float32(100))
    dn = np.nextafter(r, np.float32(0))
    z = np.float32(0)
    x = np.array([z, dn, r, up, z, z], np.float32)
    y = np.array([z, z, z, z, r, -up], np.float32)
    fc = np.array([1, 5], np.uint32)
    o = oracle_mod.Oracle.from_arrays(x, y, fc, np.array([-50, -50, 50, 50], np.float32))
    o.bin(float(r))
    rows, slots, _, _ = o.join2_count(0, 1)
    assert dn * dn < r * r and up * up > r * r
    assert rows == 3  # r - ulp, r along x, r along y; r + ulp excluded


def test_pi_compat_hand(oracle_mod):
    # features: a has 3 instances, b has 2 (degree 2); c has 1
    x = np.arange(6, dtype=np.float32)
    o = oracle_mod.Oracle.from_arrays(x, x, np.array([3, 2, 1], np.uint32), np.array([0, 0, 10, 10], np.float32))
    # degree 2, P_a = [1,0,1], P_b = [0,1]: words [1, 0, a2|b0 = 1, 1]
    bm = np.array([1, 0, 1, 1, 0], np.uint32)
    pi, num = o.pi_compat(bm, [0, 1])
    assert num.tolist() == [1, 2]  # a: words 0..1; b: words 2..3
    assert pi == np.float32(1) / np.float32(3)
    ex = np.array([1, 0, 1, 0, 1], np.uint32)
    pe, ne = o.pi_exact(ex, [0, 1])
    assert ne.tolist() == [2, 1] and pe == np.float32(0.5)
    # degree 3 over (a, b, c): P_a = [0,1,1], P_b = [1,1], P_c = [1]
    bm3 = np.array([0, 1, 1, 1, 1, 1], np.uint32)
    pi3, n3 = o.pi_compat(bm3, [0, 1, 2])
    # a: words [0,2) -> 1; b: [2,4) -> P_a[2] + P_b[0] = 2; c: [4,5) -> P_b[1] = 1
    assert n3.tolist() == [1, 2, 1]
    assert pi3 == np.float32(1) / np.float32(3)
    pe3, e3 = o.pi_exact(bm3, [0, 1, 2])
    assert e3.tolist() == [2, 2, 1]
    # N = 1 first feature: its window is empty
    pi1, n1 = o.pi_compat(np.array([1, 1, 1], np.uint32), [2, 1])
    assert n1.tolist() == [0, 2] and pi1 == 0.0


def test_filter_literal_vs_fast_and_sound(oracle_mod):
    rng = np.random.default_rng(9)
    x, y, fc, b = _dataset(rng, 300, 5, 400.0, hot=0.4)
    o = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    r = 30.0
    o.bin(r)
    pairs = np.array([[i, j] for i in range(5) for j in range(i + 1, 5)], np.uint32)
    trip = np.array([[i, j, k] for i in range(5) for j in range(i + 1, 5) for k in range(j + 1, 5)], np.uint32)
    for cands, k in ((pairs, 2), (trip, 3)):
        c1, p1 = o.filter(cands, k, 0.2)
        c2, p2 = o.filter(cands, k, 0.2, literal=True)
        np.testing.assert_array_equal(c1, c2)
        np.testing.assert_array_equal(p1, p2)
    counts, _ = o.filter(pairs, 2, 0.0)
    for (a, bb), cnt in zip(pairs.tolist(), counts):
        _, _, _, ex = o.join2_count(a, bb)
        assert cnt[0] >= ex[: fc[a]].sum() and cnt[1] >= ex[fc[a]: fc[a] + fc[bb]].sum()


def test_candgen_first_failure_break(oracle_mod):
    L = np.array([[0, 1], [0, 2], [0, 3], [1, 3], [2, 3]], np.uint32)
    # i=(0,1), j=(0,2): (0,1,2) lacks (1,2) -> the j loop ends, so (0,1,3) is
    # never formed; i=(0,2), j=(0,3): (0,2,3) has (2,3).
    assert oracle_mod.candgen(L, 2).tolist() == [[0, 2, 3]]
    L2 = np.array([[0, 1], [0, 2], [0, 3], [1, 2], [2, 3]], np.uint32)
    assert oracle_mod.candgen(L2, 2).tolist() == [[0, 1, 2], [0, 2, 3]]
    L3 = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]], np.uint32)
    assert oracle_mod.candgen(L3, 3).tolist() == [[0, 1, 2, 3]]
    L4 = np.array([[0, 1, 2], [0, 1, 3], [0, 1, 4], [0, 2, 3], [1, 2, 3], [0, 3, 4]], np.uint32)
    # (0,1,2,3) ok; (0,1,2,4) lacks (1,2,4) -> break: (0,1,3,4) never formed
    assert oracle_mod.candgen(L4, 3).tolist() == [[0, 1, 2, 3]]
    assert oracle_mod.candgen(np.zeros((0, 2), np.uint32), 2).shape[0] == 0


def _report_sets(rep):
    pats = {}
    for l in rep.splitlines():
        if l.startswith("c ") and l.endswith("v 1"):
            f = l.split(" f ")[0].split()[1:]
            pats.setdefault(len(f), set()).add(tuple(map(int, f)))
    return pats


def _report_rows(rep):
    out = {}
    for l in rep.splitlines():
        if l.startswith("c ") and " r " in l:
            f = tuple(map(int, l.split(" f ")[0].split()[1:]))
            e = list(map(int, l.split(" e ")[1].split(" pi ")[0].split()))
            out[f] = e
    return out


@pytest.mark.parametrize("seed", [0, 1])
def test_miner_filter_invariance_and_monotone(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    x, y, fc, b = _dataset(rng, 300, 6, 800.0, hot=0.3)
    o = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    r = 25.0
    rep_on = o.mine(r, 0.2, max_k=4)
    rep_off = o.mine(r, 0.2, max_k=4, filter_on=False)
    assert _report_sets(rep_on) == _report_sets(rep_off)
    assert sum(len(v) for v in _report_sets(rep_on).values()) > 0
    rows = _report_rows(rep_off)
    for f, e in rows.items():
        if len(f) >= 3 and f[:-1] in rows:
            pe = rows[f[:-1]]
            for j in range(len(f) - 1):
                assert e[j] <= pe[j], (f, e, pe)


def _f32_nearest(s):
    """Correctly rounded decimal -> binary32 (ties to even), exactly."""
    F = Fraction(s)
    if F == 0:
        return np.float32(-0.0) if s.strip().startswith("-") else np.float32(0.0)
    g = np.float32(float(F))
    cands = [np.nextafter(g, np.float32(-np.inf)), g, np.nextafter(g, np.float32(np.inf))]
    best = min(cands, key=lambda c: (abs(Fraction(float(c)) - F), int(np.float32(c).view(np.uint32)) & 1))
    return np.float32(best)


def test_csv_parse(oracle_mod, tmp_path):
    vals = ["0.1", "16011.3", "24902.3", "7719.56", "1e-3", "3.4028234663852886e+38", "0.30000001192092896",
            "1.0000000596046448", "2.5", "-0.0"]
    rows = [("7", vals[i], vals[-1 - i]) for i in range(3)] + [("7", "1", "2"), ("3", "4.5", "5.5"),
                                                               ("7", vals[7], vals[6])]
    path = tmp_path / "d.csv"
    with open(path, "w") as fh:
        fh.write(f"{len(rows)}\n")
        for r in rows:
            fh.write(",".join(r) + "\n")
        fh.write("\n0,0,16011.3,24902.3")
    o = oracle_mod.Oracle.from_csv(str(path))
    assert o.F == 3  # 7, 3, 7 again: a label that comes back is a new feature
    assert o.fcount.tolist() == [4, 1, 1]
    assert [o.label(i) for i in range(3)] == ["7", "3", "7"]
    x, y = o.points()
    ex = [_f32_nearest(r[1]) for r in rows]
    ey = [_f32_nearest(r[2]) for r in rows]
    assert x.view(np.uint32).tolist() == np.array(ex, np.float32).view(np.uint32).tolist()
    assert y.view(np.uint32).tolist() == np.array(ey, np.float32).view(np.uint32).tolist()
    assert o.bounds().tolist() == [0.0, 0.0, float(_f32_nearest("16011.3")), float(_f32_nearest("24902.3"))]


@pytest.mark.parametrize("body", ["2\n0,1,2\n0,3,4\n0,0,5,5",          # no empty line
                                  "3\n0,1,2\n0,3,4\n\n0,0,5,5",        # N mismatch
                                  "2\n0,1,2\nx,3,4\n\n0,0,5,5",        # bad label
                                  "2\n0,1,2\n0,3\n\n0,0,5,5"])          # short row
def test_csv_malformed(oracle_mod, tmp_path, body):
    path = tmp_path / "bad.csv"
    path.write_text(body)
    with pytest.raises(ValueError):
        oracle_mod.Oracle.from_csv(str(path))


def test_table_hash_properties():
    """The report's table checksum: a sum of per-row hashes (the value of an
    empty table is 0), so it ignores row order, and it changes with any
    instance id, with column order and with a duplicated row."""
    import oracle
    rng = np.random.default_rng(7)
    t = rng.integers(0, 1 << 31, size=(500, 3)).astype(np.uint32)
    h = oracle.table_hash(t)
    assert oracle.table_hash(t[rng.permutation(len(t))]) == h
    assert oracle.table_hash(t[:0]) == 0
    assert h == sum(oracle.row_hash(r) for r in t) % (1 << 64)
    u = t.copy()
    u[123, 1] ^= 1
    assert oracle.table_hash(u) != h
    assert oracle.table_hash(t[:, [1, 0, 2]]) != h
    assert oracle.table_hash(np.vstack([t, t[:1]])) != h
