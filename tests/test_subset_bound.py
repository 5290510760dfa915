"""The subset bound of the miner (no counterpart in the reference), pinned on
the oracle:

- every bound equals the one computed from the participation sets of the
  oracle's own instance tables (np.unique of each subset table's columns,
  intersected), not from its bitmaps;
- the bound is sound: it is at least the PI of every candidate it would
  skip, in both numerator modes, and mining with it leaves the prevalent
  patterns, the instance tables, the pattern files and every counted
  candidate's line unchanged.
"""
import itertools
import os

import numpy as np
import pytest

from report_check import OracleChains, fbits, parse_report

# 14 features of 20k instances around 30 shared hotspots: every pair is
# prevalent at theta 0.15 and most of the 364 triples are not
DENSE = dict(N=20000, seed=3, F=14, hot_frac=0.5, hotspots=30, sigma=(300, 300), theta=0.15, r=80.0)


@pytest.fixture(scope="module")
def dense(datagen, oracle_mod):
    labels, x, y, S, p = datagen.generate("c1", **DENSE)
    fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
    orc = oracle_mod.Oracle.from_arrays(x.astype(np.float32), y.astype(np.float32), fc,
                                        np.array([0, 0, S, S], np.float32))
    orc.bin(p["r"])
    return orc, fc, p


def _bound_from_tables(chains, fc, cand, pi_mode):
    fstart = np.concatenate([[0], np.cumsum(fc.astype(np.int64))[:-1]])
    k = len(cand)
    cnt, last = [], []
    for j, f in enumerate(cand):
        inter = None
        for S in itertools.combinations(cand, k - 1):
            if f not in S:
                continue
            col = chains.table(S).reshape(-1, k - 1)[:, S.index(f)]
            u = np.unique(col)
            inter = u if inter is None else np.intersect1d(inter, u)
        cnt.append(len(inter))
        last.append(int(fstart[f] + fc[f] - 1 in set(inter.tolist())))
    q = np.float32(1.0)
    for j, f in enumerate(cand):
        v = cnt[j] + (last[j - 1] if pi_mode == 0 and j >= 1 else 0)
        b = np.float32(v) / np.float32(fc[f])
        q = min(q, b)
    return q


@pytest.mark.parametrize("pi_mode", [0, 1])
def test_bound_equals_table_participation(dense, pi_mode):
    orc, fc, p = dense
    rep = orc.mine(p["r"], p["theta"], max_k=4, pi_mode=pi_mode, subset_bound=True)
    cands, _, _ = parse_report(rep)
    chains = OracleChains(orc)
    with_q = [c for c, d in cands.items() if "q" in d]
    assert len(with_q) > 300 and all(len(c) >= 3 for c in with_q)
    for c in with_q:
        assert fbits(_bound_from_tables(chains, fc, c, pi_mode)) == cands[c]["q"], c


def _strip_q(line):
    t = line.split()
    if "q" in t:
        i = t.index("q")
        del t[i:i + 2]
    return " ".join(t)


@pytest.mark.parametrize("pi_mode", [0, 1])
@pytest.mark.parametrize("filter_on", [True, False])
def test_bound_sound_and_invisible(dense, pi_mode, filter_on, tmp_path):
    orc, fc, p = dense
    th = np.float32(p["theta"])
    off = orc.mine(p["r"], p["theta"], max_k=4, pi_mode=pi_mode, filter_on=filter_on,
                   out_prefix=str(tmp_path / "off_"))
    on = orc.mine(p["r"], p["theta"], max_k=4, pi_mode=pi_mode, filter_on=filter_on, subset_bound=True,
                  out_prefix=str(tmp_path / "on_"))
    c_off, t_off, l_off = parse_report(off)
    c_on, t_on, l_on = parse_report(on)
    assert t_on == t_off and l_on == l_off
    lines_off = {tuple(l.split(" f ")[0].split()[1:]): l for l in off.splitlines() if l.startswith("c ")}
    pruned = 0
    for line in on.splitlines():
        if not line.startswith("c "):
            continue
        key = tuple(line.split(" f ")[0].split()[1:])
        cand = tuple(int(v) for v in key)
        d_on, d_off = c_on[cand], c_off[cand]
        if "rows" in d_on:
            assert _strip_q(line) == lines_off[key]
        else:
            pruned += 1
            assert d_off["v"] == 0, cand
        if "q" in d_on:
            q = np.uint32(int(d_on["q"], 16)).view(np.float32)
            pi = np.uint32(int(d_off["pie" if pi_mode else "pic"], 16)).view(np.float32)
            assert pi <= q, (cand, pi, q)
            assert (q < th) == ("rows" not in d_on)
    assert pruned > 200
    for k in (2, 3, 4):
        a, b = tmp_path / f"off_ColocationDegree_{k}.txt", tmp_path / f"on_ColocationDegree_{k}.txt"
        assert os.path.exists(a) == os.path.exists(b)
        if os.path.exists(a):
            assert a.read_bytes() == b.read_bytes()
