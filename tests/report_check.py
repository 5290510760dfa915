"""Helpers of the parity tests: parsing a colo-report v1 text and checking a
sample of its candidates against the oracle.

A report (written identically by the GPU miner and by the oracle's orc_mine)
has one line per candidate:
    c <ids> f <filter counts> p <pass> [q <subset bound bits>] [r <rows>
    n <compat num> e <exact num> pi <compat PI bits> <exact PI bits>
    v <prevalent>]
(q only when mining with the subset bound, from size 3 on; a candidate
whose bound is below theta has no counts)
and one per instance table the reference would build:
    t <ids> rows <rows> [h <row-order independent checksum>]
"""
import hashlib

import numpy as np


def parse_report(text):
    """Returns ({cand: fields}, {cand: (rows, checksum or None)}, [level lines])."""
    cands, tables, levels = {}, {}, []
    for line in text.splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "c":
            fi = t.index("f")
            k = fi - 1
            cand = tuple(int(v) for v in t[1:fi])
            d = {"f": [int(v) for v in t[fi + 1:fi + 1 + k]], "p": int(t[t.index("p") + 1])}
            if "q" in t:
                d["q"] = t[t.index("q") + 1]
            if "r" in t:
                d["rows"] = int(t[t.index("r") + 1])
                ni, ei, pi = t.index("n"), t.index("e"), t.index("pi")
                d["numc"] = [int(v) for v in t[ni + 1:ni + 1 + k]]
                d["nume"] = [int(v) for v in t[ei + 1:ei + 1 + k]]
                d["pic"], d["pie"] = t[pi + 1], t[pi + 2]
                d["v"] = int(t[t.index("v") + 1])
            cands[cand] = d
        elif t[0] == "t":
            ri = t.index("rows")
            cand = tuple(int(v) for v in t[1:ri])
            tables[cand] = (int(t[ri + 1]), t[t.index("h") + 1] if "h" in t else None)
        elif t[0] == "level":
            levels.append(line)
    return cands, tables, levels


def fbits(v):
    return "%08x" % int(np.float32(v).view(np.uint32))


def dataset_fingerprint(x, y, fcount):
    h = hashlib.sha256()
    for a in (np.ascontiguousarray(x, np.float32), np.ascontiguousarray(y, np.float32),
              np.ascontiguousarray(fcount, np.uint32)):
        h.update(a.tobytes())
    return h.hexdigest()


def choose_sample(text, seed, per_level=None, max_parent_rows=3e7):
    """Seeded sample of a report's candidates: at level 2 the five with the
    most rows plus random ones; at every level random counted candidates
    whose prefix tables hold at most max_parent_rows rows (the oracle
    rebuilds the prefix chain), plus up to two filtered-out candidates."""
    per_level = per_level or {2: 15, 3: 10, 4: 5, 5: 3, 6: 3}
    cands, _, _ = parse_report(text)
    rng = np.random.default_rng(seed)
    rows = {c: d["rows"] for c, d in cands.items() if "rows" in d}
    out = []
    for k in sorted({len(c) for c in cands}):
        counted = sorted(c for c in cands if len(c) == k and "rows" in cands[c])
        if k == 2:
            top = sorted(counted, key=lambda c: -rows[c])[:5]
            out.extend(top)
            counted = [c for c in counted if c not in top]
        else:
            counted = [c for c in counted
                       if all(rows.get(c[:j], 1e30) <= max_parent_rows for j in range(2, k))]
        n = min(per_level.get(k, 3), len(counted))
        if n:
            out.extend(counted[i] for i in sorted(rng.choice(len(counted), n, replace=False)))
        cut = sorted(c for c in cands if len(c) == k and not cands[c]["p"])
        if cut:
            out.extend(cut[i] for i in sorted(rng.choice(len(cut), min(2, len(cut)), replace=False)))
    return out


class OracleChains:
    """Prefix tables built by the oracle's own write passes (join2_write,
    joink_write), cached: the parents of sampled level-k candidates."""

    def __init__(self, orc):
        self.orc = orc
        self.t = {}

    def table(self, pat):
        pat = tuple(pat)
        if pat in self.t:
            return self.t[pat]
        if len(pat) == 2:
            _, slots, _, _ = self.orc.join2_count(*pat)
            tab = self.orc.join2_write(pat[0], pat[1], slots)
        else:
            parent = self.table(pat[:-1])
            _, slots, _ = self.orc.joink_count(parent, len(pat), list(pat))
            tab = self.orc.joink_write(parent, len(pat), list(pat), slots)
        self.t[pat] = tab
        return tab


def check_sample(oracle_mod, orc, text, theta, sample, say=print):
    """Every sampled candidate of the report against the oracle: filter
    counts and pass flag, rows, compat and exact numerators, both PIs as
    binary32 bits, the prevalence verdict, and the checksum of its instance
    table when the report lists one.  Returns the number of candidates and
    tables checked."""
    cands, tables, _ = parse_report(text)
    chains = OracleChains(orc)
    by_k = {}
    for c in sample:
        by_k.setdefault(len(c), []).append(c)
    n_c = n_t = 0
    for k, cs in sorted(by_k.items()):
        oc, op = orc.filter(np.array(cs, np.uint32), k, theta)
        for c, cnt, ps in zip(cs, oc, op):
            d = cands[c]
            assert [int(v) for v in cnt] == d["f"], (c, list(cnt), d["f"])
            assert int(ps) == d["p"], c
        for c in cs:
            d = cands[c]
            if "rows" not in d:
                continue
            say(f"oracle check {c}")
            if k == 2:
                rows, slots, bm, ex = orc.join2_count(*c)
                pic, numc = orc.pi_compat(bm, list(c))
                pie, nume = orc.pi_exact(ex, list(c))
            else:
                parent = chains.table(c[:-1])
                rows, slots, bm = orc.joink_count(parent, k, list(c))
                pic, numc = orc.pi_compat(bm, list(c))
                pie, nume = orc.pi_exact(bm, list(c))
            assert rows == d["rows"], (c, rows, d["rows"])
            assert [int(v) for v in numc] == d["numc"], (c, list(numc), d["numc"])
            assert [int(v) for v in nume] == d["nume"], (c, list(nume), d["nume"])
            assert fbits(pic) == d["pic"] and fbits(pie) == d["pie"], (c, fbits(pic), fbits(pie), d)
            pim = pic  # compat numerators decide prevalence (pi_mode 0)
            assert int(not (np.float32(pim) < np.float32(theta))) == d["v"], c
            n_c += 1
            if c in tables and tables[c][1] is not None:
                tab = chains.table(c) if k == 2 or c[:-1] in chains.t else None
                if tab is None:
                    tab = orc.joink_write(chains.table(c[:-1]), k, list(c), slots)
                assert tables[c][0] == len(tab) // k, c
                h = "%016x" % oracle_mod.table_hash(tab.reshape(-1, k))
                assert h == tables[c][1], (c, h, tables[c][1])
                n_t += 1
    return n_c, n_t
