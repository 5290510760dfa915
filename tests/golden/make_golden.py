#!/usr/bin/env python3
"""Writes the oracle's full mining reports of generated workloads, and of
the reference's own TestCase (Seattle 2012), as golden fixtures for the GPU
parity tests (tests/test_bench_configs.py, tests/test_gpu_parity.py).

For each workload the generator (gpu-colocationmininig_amd/datagen.py) is
run with its fixed seed, the oracle (oracle/colo_oracle.c, the CPU
restatement of the reference) mines it completely with instance-table
checksums, and tests/golden/<name>_report.txt.gz receives the report
preceded by one header line:
    # workload <name> theta <theta> max_k <k> sha256 <fingerprint of x, y, fcount>
The fingerprint lets a test detect a generator change before comparing.

"seattle" mines tests/golden/seattle2012.csv.gz (the reference's
TestCase/Data/Seattle2012.csv) with the TestCase parameters r = 10,
theta = 0.3 (TestCase/Data/parameter.txt:1-2) and no size cap, and also
writes the pattern files as tests/golden/seattle_ColocationDegree_<k>.txt.

    python tests/golden/make_golden.py c3 seattle [--threads 8]
"""
import argparse
import glob
import gzip
import os
import shutil
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "gpu-colocationmininig_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import datagen  # noqa: E402
import oracle  # noqa: E402
from report_check import dataset_fingerprint  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("names", nargs="+")
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    if a.threads:
        oracle.set_threads(a.threads)
    for name in a.names:
        t0 = time.time()
        if name == "seattle":
            seattle(t0)
            continue
        labels, x, y, S, p = datagen.generate(name)
        fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
        x32, y32 = x.astype(np.float32), y.astype(np.float32)
        del labels, x, y
        fp = dataset_fingerprint(x32, y32, fc)
        orc = oracle.Oracle.from_arrays(x32, y32, fc, np.array([0, 0, S, S], np.float32))
        rep = orc.mine(p["r"], p["theta"], max_k=p["max_k"], hash_tables=True)
        out = os.path.join(HERE, f"{name}_report.txt.gz")
        with gzip.open(out, "wt") as fh:
            fh.write(f"# workload {name} theta {p['theta']!r} max_k {p['max_k']} sha256 {fp}\n")
            fh.write(rep)
        print(f"{name}: {time.time() - t0:.0f} s, {rep.splitlines()[-1]} -> {out}", flush=True)


def seattle(t0):
    d = tempfile.mkdtemp(prefix="seattle_golden_")
    csv = os.path.join(d, "Seattle2012.csv")
    with gzip.open(os.path.join(HERE, "seattle2012.csv.gz"), "rb") as fi, open(csv, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    orc = oracle.Oracle.from_csv(csv)
    x, y = orc.points()
    fp = dataset_fingerprint(x, y, orc.fcount)
    rep = orc.mine(10.0, 0.3, max_k=0, hash_tables=True, out_prefix=os.path.join(d, ""))
    with gzip.open(os.path.join(HERE, "seattle_report.txt.gz"), "wt") as fh:
        fh.write(f"# workload seattle theta 0.3 max_k 0 sha256 {fp}\n")
        fh.write(rep)
    for f in sorted(glob.glob(os.path.join(d, "ColocationDegree_*.txt"))):
        shutil.copy(f, os.path.join(HERE, "seattle_" + os.path.basename(f)))
    shutil.rmtree(d)
    print(f"seattle: {time.time() - t0:.0f} s, {rep.splitlines()[-1]}", flush=True)


if __name__ == "__main__":
    main()
