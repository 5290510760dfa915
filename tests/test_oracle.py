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
    # exact numerators never exceed the filter's bound, so the filter cannot
    # change the exact-mode pattern set; a compat numerator (off-by-one
    # windows) can exceed the exact one by 1 and so can exceed a tight bound
    rep_on = o.mine(r, 0.2, max_k=4, pi_mode=1)
    rep_off = o.mine(r, 0.2, max_k=4, pi_mode=1, filter_on=False)
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


def filter_bruteforce(x, y, fc, b, r, cands, theta):
    """The cell-aggregate bound (K:469-509, 545-573) restated on arrays,
    independently of the oracle's code: per-(cell, feature) counts on the
    gy x gx grid; a 2x2 quadrant qualifies when every member occurs in one
    of its four cells; every cell of the union of qualifying quadrants counts
    once; ub_t = fl32(double(cnt_t) / double(fl32(N_t))); pass iff
    min_t ub_t >= theta.  Returns (counts, pass, min_t ub_t)."""
    cid, gx = _cells32(x, y, b, r)
    r32 = np.float32(r)
    gy = int(np.ceil(np.float32(np.float32(b[3]) + r32 - np.float32(b[1])) / r32))
    F = len(fc)
    fid = np.repeat(np.arange(F), fc)
    cnt = np.zeros((gy * gx, F), np.int64)
    np.add.at(cnt, (cid.astype(np.int64), fid), 1)
    cnt = cnt.reshape(gy, gx, F)
    out_c, out_p, out_u = [], [], []
    for cand in cands:
        q = np.ones((max(gy - 1, 0), max(gx - 1, 0)), bool)
        for f in cand:
            pr = cnt[:, :, f] > 0
            q &= pr[:-1, :-1] | pr[1:, :-1] | pr[:-1, 1:] | pr[1:, 1:]
        U = np.zeros((gy, gx), bool)
        U[:-1, :-1] |= q
        U[1:, :-1] |= q
        U[:-1, 1:] |= q
        U[1:, 1:] |= q
        c = [int(cnt[:, :, f][U].sum()) for f in cand]
        ub = min(np.float32(np.float64(ci) / np.float64(np.float32(fc[f]))) for ci, f in zip(c, cand))
        out_c.append(c)
        out_p.append(bool(ub >= np.float32(theta)))
        out_u.append(ub)
    return np.array(out_c, np.uint32), np.array(out_p), out_u


@pytest.mark.parametrize("seed", range(40))
def test_filter_bruteforce(oracle_mod, seed):
    """The oracle's filter counts and pass flags equal the array restatement
    on small random grids (clustered and uniform points, pairs and triples),
    with theta also placed exactly on a bound (passes) and one float32 step
    above it (fails)."""
    rng = np.random.default_rng(1000 + seed)
    F = int(rng.integers(3, 6))
    S = float(rng.choice([100.0, 150.0, 250.0]))
    r = float(rng.choice([20.0, 30.0, 45.0]))
    x, y, fc, b = _dataset(rng, int(rng.integers(15, 60)), F, S, hot=float(rng.choice([0.0, 0.5])))
    o = oracle_mod.Oracle.from_arrays(x, y, fc, b)
    o.bin(r)
    for k in (2, 3):
        cands = np.array([c for c in __import__("itertools").combinations(range(F), k)], np.uint32)
        ec, ep, eu = filter_bruteforce(x, y, fc, b, r, cands, 0.3)
        oc, op = o.filter(cands, k, 0.3)
        np.testing.assert_array_equal(oc, ec)
        np.testing.assert_array_equal(op, ep)
        lit_c, lit_p = o.filter(cands, k, 0.3, literal=True)
        np.testing.assert_array_equal(lit_c, ec)
        u = max(eu)  # theta exactly on a bound, and one step above it
        if u > 0:
            for th, want in ((u, True), (np.nextafter(np.float32(u), np.float32(2)), False)):
                _, p_at = o.filter(cands, k, float(th))
                _, e_at, _ = filter_bruteforce(x, y, fc, b, r, cands, th)
                np.testing.assert_array_equal(p_at, e_at)
                assert p_at[int(np.argmax(eu))] == want


HAND_CSV = """10
7,10.0,10.0
7,50.0,50.0
3,11.0,10.0
3,51.0,50.0
7,10.0,11.0
7,50.0,51.0
12,90.0,90.0
12,90.5,10.0
5,10.5,10.5
5,95.0,95.0

0,0,100,100"""

# Worked by hand from the reference's rules for r = 2, theta = 0.5.  Features
# in row order: f0 "7", f1 "3", f2 "7" (a reappearing label is a new feature,
# cpp:168-183), f3 "12", f4 "5".  f0, f1, f2 meet at (10, 10) and (50, 50),
# f4 only at (10, 10), f3 nowhere.  With two instances per feature the
# reference's windows (cu:116-124) give the first member 1 of 2 and every
# other member 1 or 2 of 2, so each co-located set has PI 0.5 and is
# prevalent (cpp:335: !(PI < theta)); sets with f3 have PI 0.  Level 3 from
# the level-2 list by the Apriori join with its first-failure break
# (cpp:517-557): (0,1,2), (0,1,4), (0,2,4), (1,2,4); level 4: (0,1,2,4).
# Lines per savetoFile2 / savetoFileGen (cpp:960-997).
HAND_FILES = {
    2: "( 7 | 3 )\n( 7 | 7 )\n( 7 | 5 )\n( 3 | 7 )\n( 3 | 5 )\n( 7 | 5 )\n",
    3: "( 7 | 3 | 7 )\n( 7 | 3 | 5 )\n( 7 | 7 | 5 )\n( 3 | 7 | 5 )\n",
    4: "( 7 | 3 | 7 | 5 )\n",
}


def test_pattern_files_hand(oracle_mod, tmp_path):
    """The oracle's ColocationDegree_<k>.txt files equal the hand-worked
    expectation byte for byte, and nothing is written past size 4."""
    csv = tmp_path / "hand.csv"
    csv.write_text(HAND_CSV)
    o = oracle_mod.Oracle.from_csv(str(csv))
    assert [o.label(f) for f in range(o.F)] == ["7", "3", "7", "12", "5"]
    prefix = str(tmp_path) + "/"
    rep = o.mine(2.0, 0.5, max_k=0, out_prefix=prefix)
    for k, text in HAND_FILES.items():
        assert (tmp_path / f"ColocationDegree_{k}.txt").read_text() == text, k
    assert not (tmp_path / "ColocationDegree_5.txt").exists()
    assert rep.splitlines()[-1] == "end levels 3 patterns 11"


def _round_f32(fr):
    """Fraction -> nearest binary32 (ties to even), exactly."""
    c = np.float32(float(fr))
    best = None
    for v in (np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))):
        d = abs(Fraction(float(v)) - fr)
        if best is None or d < best[0] or (d == best[0] and int(v.view(np.uint32)) % 2 == 0):
            best = (d, v)
    return best[1]


def test_boundary_pair_forms(oracle_mod):
    """orc_boundary_pairs against exact rational arithmetic: pairs placed
    within a few ulps of r^2 (far apart from each other, so each instance's
    3x3 cells hold only its partner); for each pair the canonical form
    fl(fl(dx*dx) + fl(dy*dy)), both fused forms (exact product plus the
    other rounded square, rounded once) and the double-accumulated form
    (exact sum rounded to binary64, then to binary32) are evaluated here and
    their <= r^2 decisions compared."""
    rng = np.random.default_rng(21)
    r = np.float32(7.3)
    r2 = r * r
    xs, ys, qx, qy = [], [], [], []
    while len(xs) < 3000:
        k = len(xs)
        ax, ay = np.float32(50.0 * (k % 60) + 11.0), np.float32(50.0 * (k // 60) + 11.0)
        t = rng.uniform(0, 2 * np.pi)
        d = float(r) * (1 + rng.uniform(-4e-7, 4e-7))
        bx, by = np.float32(float(ax) + d * np.cos(t)), np.float32(float(ay) + d * np.sin(t))
        dx, dy = bx - ax, by - ay
        if abs(float(dx * dx + dy * dy) - float(r2)) > 16 * float(np.spacing(r2)):
            continue
        xs.append(ax); ys.append(ay); qx.append(bx); qy.append(by)
    n = len(xs)
    x = np.array(xs + qx, np.float32)
    y = np.array(ys + qy, np.float32)
    S = np.float32(50.0 * 61)
    o = oracle_mod.Oracle.from_arrays(x, y, np.array([n, n], np.uint32), np.array([0, 0, S, S], np.float32))
    o.bin(float(r))
    got = o.boundary_pairs()
    want = dict(tests=n, neighbours=0, fma_dx_first=0, fma_dy_first=0, double_sum=0, any_form=0)
    R2 = Fraction(float(r2))
    for i in range(n):
        dx, dy = np.float32(qx[i]) - np.float32(xs[i]), np.float32(qy[i]) - np.float32(ys[i])
        sx, sy = dx * dx, dy * dy
        canon = bool(sx + sy <= r2)
        fa = Fraction(float(dx)) ** 2 + Fraction(float(sy))
        fb = Fraction(float(dy)) ** 2 + Fraction(float(sx))
        ra = bool(_round_f32(fa) <= r2)
        rb = bool(_round_f32(fb) <= r2)
        rd = bool(np.float32(float(Fraction(float(dx)) ** 2 + Fraction(float(dy)) ** 2)) <= r2)
        want["neighbours"] += canon
        want["fma_dx_first"] += ra != canon
        want["fma_dy_first"] += rb != canon
        want["double_sum"] += rd != canon
        want["any_form"] += ra != canon or rb != canon or rd != canon
    assert got == want
    assert want["any_form"] > 0 and want["neighbours"] > 0  # the boundary band is exercised
