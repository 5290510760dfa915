"""CPU checks of the product library (no GPU calls): it loads, exports every
entry point include/colo.h declares, and its host-side parts (CSV parser,
candidate generation) agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^[A-Za-z_][\w \*]*?\b(colo_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_exports_every_declared_symbol(colo_mod):
    L = colo_mod.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(colo_mod.SYMBOLS) | {"colo_candidates"}


def test_csv_parser_matches_oracle(colo_mod, oracle_mod, datagen, tmp_path):
    w = datagen.write_workload(str(tmp_path), "c1", N=3000)
    hd = colo_mod.HostData(w["csv"])
    o = oracle_mod.Oracle.from_csv(w["csv"])
    x, y = o.points()
    assert hd.n == o.n and hd.F == o.F
    assert hd.x.view(np.uint32).tolist() == x.view(np.uint32).tolist()
    assert hd.y.view(np.uint32).tolist() == y.view(np.uint32).tolist()
    assert hd.fcount.tolist() == o.fcount.tolist()
    assert hd.bounds.tolist() == o.bounds().tolist()
    assert hd.labels == [o.label(f) for f in range(o.F)]


def test_csv_parser_seattle(colo_mod, oracle_mod, tmp_path):
    import gzip
    src = os.path.join(ROOT, "tests", "golden", "seattle2012.csv.gz")
    csv = tmp_path / "s.csv"
    with gzip.open(src, "rb") as fi:
        csv.write_bytes(fi.read())
    hd = colo_mod.HostData(str(csv))
    o = oracle_mod.Oracle.from_csv(str(csv))
    x, y = o.points()
    assert hd.n == 165458 and hd.F == 13
    assert hd.x.view(np.uint32).tolist() == x.view(np.uint32).tolist()
    assert hd.y.view(np.uint32).tolist() == y.view(np.uint32).tolist()
    assert hd.fcount.tolist() == o.fcount.tolist()


@pytest.mark.parametrize("body", ["2\n0,1,2\n0,3,4\n0,0,5,5", "3\n0,1,2\n0,3,4\n\n0,0,5,5",
                                  "2\n0,1,2\nx,3,4\n\n0,0,5,5"])
def test_csv_parser_rejects(colo_mod, tmp_path, body):
    p = tmp_path / "bad.csv"
    p.write_text(body)
    with pytest.raises(colo_mod.ColoError):
        colo_mod.HostData(str(p))


def test_candidates_match_oracle(colo_mod, oracle_mod):
    rng = np.random.default_rng(4)
    for km1 in (2, 3, 4):
        for trial in range(30):
            F = int(rng.integers(km1 + 1, 12))
            allp = [tuple(sorted(rng.choice(F, km1, replace=False).tolist())) for _ in range(60)]
            L = np.array(sorted(set(allp)), np.uint32)
            np.testing.assert_array_equal(colo_mod.candidates(L, km1), oracle_mod.candgen(L, km1))
    L = np.array([[0, 1], [0, 2], [0, 3], [1, 3], [2, 3]], np.uint32)
    assert colo_mod.candidates(L, 2).tolist() == [[0, 2, 3]]


def test_no_gpu_fails_loudly(colo_mod):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(colo_mod.ColoError):
        colo_mod.Context(np.zeros(4, np.float32), np.zeros(4, np.float32), [2, 2], [0, 0, 1, 1], 0.5)


def test_compat_header_compiles(tmp_path):
    """include/DeviceProcess_compat.h keeps the reference's entry-point names
    (Codes/DeviceProcess.h:4-36) and must compile as a C++ header."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = tmp_path / "use.cpp"
    src.write_text('#include "DeviceProcess_compat.h"\nusing namespace colo_compat;\n'
                   "int main() { Integer *s = nullptr; (void)degree2InstanceTableSlotCounter(0, s, 0, 0, 0, s, 0);\n"
                   "  (void)scalebyConstant(2, nullptr, 0); return (int)getTotalInstances(s, 0); }\n")
    r = subprocess.run([hipcc, "-std=c++17", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"), str(src)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


def _hd_equal(a, b):
    assert (a.n, a.F) == (b.n, b.F)
    assert a.x.view(np.uint32).tolist() == b.x.view(np.uint32).tolist()
    assert a.y.view(np.uint32).tolist() == b.y.view(np.uint32).tolist()
    assert a.fcount.tolist() == b.fcount.tolist()
    assert a.labels == b.labels
    assert a.bounds.view(np.uint32).tolist() == b.bounds.view(np.uint32).tolist()


def test_csv_parallel_equals_sequential(colo_mod, datagen, tmp_path):
    """The multi-threaded parse equals the one-thread parse on a generated
    dataset (ranges split mid-feature) and on small files whose thread
    ranges cut labels, quoted fields and repeated labels at every offset."""
    w = datagen.write_workload(str(tmp_path / "w"), "c2", N=120_000)
    one = colo_mod.HostData(w["csv"], threads=1)
    for t in (2, 7, 16):
        _hd_equal(colo_mod.HostData(w["csv"], threads=t), one)
    _hd_equal(colo_mod.HostData(w["csv"]), one)
    rng = np.random.default_rng(3)
    for trial in range(20):
        n = int(rng.integers(1, 40))
        labs = np.sort(rng.integers(0, 6, n))
        if trial % 3 == 0:  # a label that reappears later becomes a new feature
            labs[n // 2:] = labs[: n - n // 2]
        rows = []
        for l in labs:
            x, y = (float(v) for v in rng.uniform(0, 100, 2))
            rows.append(f'"{l}",{x!r},"{y!r}"' if rng.random() < 0.3 else f"{l},{x!r},{y!r}")
        p = tmp_path / f"s{trial}.csv"
        p.write_text(f"{n}\n" + "\n".join(rows) + "\n\n0,0,100,100")
        ref = colo_mod.HostData(str(p), threads=1)
        assert ref.fcount.sum() == n
        for t in (2, 3, 5, 8, 13):
            _hd_equal(colo_mod.HostData(str(p), threads=t), ref)


@pytest.mark.parametrize("threads", [1, 4])
@pytest.mark.parametrize("body", ["2\n0,1,2\n0,3,4\n0,0,5,5", "3\n0,1,2\n0,3,4\n\n0,0,5,5",
                                  "2\n0,1,2\nx,3,4\n\n0,0,5,5", "2\n0,1,2\n0,3,4\n\n0,0,5", "1\n0,1\n\n0,0,5,5"])
def test_csv_parallel_rejects(colo_mod, tmp_path, threads, body):
    p = tmp_path / "bad.csv"
    p.write_text(body)
    with pytest.raises(colo_mod.ColoError):
        colo_mod.HostData(str(p), threads=threads)


def test_csv_binary_cache(colo_mod, datagen, tmp_path):
    """The cache is written on the first load, read on the next, and ignored
    once the CSV changes."""
    w = datagen.write_workload(str(tmp_path / "w"), "c1", N=5000)
    cache = str(tmp_path / "data.colobin")
    a = colo_mod.HostData(w["csv"], cache=cache)
    assert os.path.getsize(cache) > 0
    _hd_equal(colo_mod.HostData(w["csv"], cache=cache), a)
    _hd_equal(colo_mod.HostData(w["csv"]), a)
    text = open(w["csv"]).read().splitlines()
    text[1] = text[1].split(",")[0] + ",1.5,2.5"
    with open(w["csv"], "w") as fh:
        fh.write("\n".join(text) + "\n")
    os.utime(w["csv"], ns=(1, 2))
    b = colo_mod.HostData(w["csv"], cache=cache)
    assert b.x[0] == np.float32(1.5) and b.y[0] == np.float32(2.5)
    _hd_equal(colo_mod.HostData(w["csv"], cache=cache), b)


def test_ctypes_mirrors_match_header(colo_mod, tmp_path):
    """The ctypes structures of the Python binding have the sizes and field
    offsets of the C structures in include/colo.h (compiled with gcc)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if not gcc:
        pytest.skip("gcc not available")
    mirrors = {"colo_dataset": colo_mod.Dataset, "colo_grid": colo_mod.Grid, "colo_mine_opts": colo_mod.MineOpts,
               "colo_tune": colo_mod.Tune, "colo_mine_stats": colo_mod.MineStats}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "colo.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run([gcc, "-std=c11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            cname, field, val = line.split()
            got[(cname, field)] = int(val)
    for cname, py in mirrors.items():
        assert got[(cname, "sizeof")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
