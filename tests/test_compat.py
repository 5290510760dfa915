"""The header-only shim include/DeviceProcess_compat.h at run time: a
colocationFinder-style driver (tests/compat/compat_driver.cpp) calls the
reference's entry-point names in the reference's level-2 order
(Codes/colocationFinder.cpp:256-425) and level-3 order (cpp:736-909); its
filter flags, candidates, PIs (binary32 bit patterns) and instance tables
(row order included) must equal the oracle's."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "compat", "compat_driver")


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _f32_bits(v):
    return int(np.float32(v).view(np.uint32))


def _parse(path):
    out = {"F2": {}, "P2": {}, "T2": {}, "F3": [], "P3": {}, "T3": {}}
    for line in open(path):
        f = line.split()
        if f[0] == "F2":
            out["F2"][(int(f[1]), int(f[2]))] = int(f[3])
        elif f[0] == "P2":
            out["P2"][(int(f[1]), int(f[2]))] = int(f[3], 16)
        elif f[0] == "T2":
            out["T2"][(int(f[1]), int(f[2]))] = int(f[3])
        elif f[0] == "F3":
            out["F3"].append(((int(f[1]), int(f[2]), int(f[3])), int(f[4])))
        elif f[0] == "P3":
            out["P3"][(int(f[1]), int(f[2]), int(f[3]))] = int(f[4], 16)
        elif f[0] == "T3":
            out["T3"][(int(f[1]), int(f[2]), int(f[3]))] = int(f[4])
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("N,seed", [(6000, None), (20000, 11)])
def test_compat_driver_matches_oracle(datagen, oracle_mod, tmp_path, N, seed):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/compat/compat_driver is not built (make -C tests/compat)")
    _say(f"compat driver c1 N={N} seed={seed}")
    w = datagen.write_workload(str(tmp_path / "case"), "c1", N=N, seed=seed)
    # binary32 values, passed as decimals that parse back to them exactly
    r, theta = float(np.float32(w["params"]["r"])), float(np.float32(w["params"]["theta"]))
    out = tmp_path / "out"
    out.mkdir()
    res = subprocess.run([DRIVER, w["csv"], repr(r), repr(theta), str(out)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    got = _parse(out / "result.txt")

    orc = oracle_mod.Oracle.from_csv(w["csv"])
    orc.bin(r)
    F = orc.F
    th = np.float32(theta)
    # level 2: every pair's filter flag; PI of the passing pairs; tables of
    # the pairs that level-3 candidates extend
    cand2 = np.array([(a, b) for a in range(F) for b in range(a + 1, F)], np.uint32)
    _, passed = orc.filter(cand2, 2, theta)
    assert got["F2"] == {(int(a), int(b)): int(p) for (a, b), p in zip(cand2, passed)}
    prev2, slots2 = [], {}
    for (a, b), p in zip(cand2.tolist(), passed):
        if not p:
            assert (a, b) not in got["P2"]
            continue
        _, slots, bm, _ = orc.join2_count(a, b)
        pi, _ = orc.pi_compat(bm, [a, b])
        assert got["P2"][(a, b)] == _f32_bits(pi), (a, b)
        if not np.float32(pi) < th:
            prev2.append((a, b))
            slots2[(a, b)] = slots
    cand3 = [tuple(int(v) for v in c) for c in oracle_mod.candgen(prev2, 2)]
    assert [c for c, _ in got["F3"]] == cand3
    prefixes = {tuple(c[:2]) for c in cand3}
    assert set(got["T2"]) == prefixes
    tables2 = {}
    for (a, b) in prefixes:
        t = orc.join2_write(a, b, slots2[(a, b)])
        tables2[(a, b)] = t
        assert got["T2"][(a, b)] == len(t) // 2
        g = np.fromfile(out / f"t2_{a}_{b}.bin", dtype=np.uint32)
        assert np.array_equal(g, t), (a, b)
    # level 3
    if cand3:
        _, passed3 = orc.filter(np.array(cand3, np.uint32), 3, theta)
        assert [p for _, p in got["F3"]] == [int(p) for p in passed3]
        n_tables = 0
        for c, p in zip(cand3, passed3):
            if not p:
                continue
            parent = tables2[c[:2]]
            _, slots, bm = orc.joink_count(parent, 3, list(c))
            pi, _ = orc.pi_compat(bm, list(c))
            assert got["P3"][c] == _f32_bits(pi), c
            if np.float32(pi) < th:
                assert c not in got["T3"]
                continue
            t = orc.joink_write(parent, 3, list(c), slots)
            assert got["T3"][c] == len(t) // 3
            g = np.fromfile(out / f"t3_{c[0]}_{c[1]}_{c[2]}.bin", dtype=np.uint32)
            assert np.array_equal(g, t), c
            n_tables += 1
        print(f"[compat] {len(cand2)} pairs, {len(prev2)} prevalent, {len(cand3)} triples, {n_tables} level-3 tables",
              flush=True)
        assert n_tables >= 1
