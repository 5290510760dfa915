"""CPU tests of the report checker used by the GPU parity tests
(tests/report_check.py): it accepts the oracle's own report and rejects a
report with one planted mistake in rows, a numerator, a PI bit pattern, a
filter count or a table checksum."""
import numpy as np
import pytest

from report_check import check_sample, choose_sample, parse_report


@pytest.fixture(scope="module")
def c1_case(datagen, oracle_mod):
    labels, x, y, S, p = datagen.generate("c1")
    fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
    orc = oracle_mod.Oracle.from_arrays(x.astype(np.float32), y.astype(np.float32), fc,
                                        np.array([0, 0, S, S], np.float32))
    rep = orc.mine(p["r"], p["theta"], max_k=p["max_k"])
    orc.bin(p["r"])
    return orc, rep, p


def test_checker_accepts_oracle_report(c1_case, oracle_mod):
    orc, rep, p = c1_case
    sample = choose_sample(rep, 1)
    n_c, n_t = check_sample(oracle_mod, orc, rep, p["theta"], sample, say=lambda m: None)
    assert n_c >= 10 and n_t >= 1


def _plant(rep, cand, field, k):
    """The report with one token of cand's line changed."""
    out = []
    for line in rep.splitlines():
        t = line.split()
        if t and t[0] in ("c", "t") and tuple(int(v) for v in t[1:1 + k]) == cand and \
                (len(t) > 1 + k and t[1 + k] in ("f", "rows")):
            if field in t and (t[0] == "c") == (field != "h"):
                i = t.index(field) + 1
                t[i] = ("%016x" % (int(t[i], 16) ^ 1)) if field == "h" else \
                       ("%08x" % (int(t[i], 16) + 1)) if field == "pi" else str(int(t[i]) + 1)
                line = " ".join(t)
        out.append(line)
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("field", ["r", "n", "e", "pi", "f", "h"])
def test_checker_rejects_planted_mistake(c1_case, oracle_mod, field):
    orc, rep, p = c1_case
    cands, tables, _ = parse_report(rep)
    pool = sorted(tables) if field == "h" else sorted(c for c, d in cands.items() if d["p"])
    cand = pool[0]
    bad = _plant(rep, cand, field, len(cand))
    assert bad != rep
    with pytest.raises(AssertionError):
        check_sample(oracle_mod, orc, bad, p["theta"], [cand], say=lambda m: None)
