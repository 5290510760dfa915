#!/usr/bin/env python3
"""Boundary-pair count per workload (test infrastructure, CPU only).

For every distance test of the reference's level-2 scan (each instance
against every instance of a later feature in its 3x3 cells,
Codes/DeviceProcess_Kernel.h:24-41), counts the tests where a fused
multiply-add form or a double-accumulated form of the squared distance
(Codes/Utility_Inline.h:31, pow(x2-x1,2) + pow(y2-y1,2)) decides
d2 <= r^2 differently from the canonical binary32 form that the oracle and
the HIP kernels use (orc_boundary_pairs in colo_oracle.c).  How nvcc
evaluated that line is the one rounding question the reference's source
leaves open; these counts bound how many tests it could affect.

    python oracle/boundary_pairs.py c1 c2 c3 c5 seattle [--threads N]
Prints one JSON line per workload.
"""
import argparse
import gzip
import json
import os
import shutil
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "gpu-colocationmininig_amd"))

import numpy as np  # noqa: E402

import oracle  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("names", nargs="+")
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    if a.threads:
        oracle.set_threads(a.threads)
    for name in a.names:
        t0 = time.time()
        if name == "seattle":  # the reference's TestCase at its parameters (r = 10)
            d = tempfile.mkdtemp()
            csv = os.path.join(d, "s.csv")
            with gzip.open(os.path.join(ROOT, "tests", "golden", "seattle2012.csv.gz"), "rb") as fi, \
                    open(csv, "wb") as fo:
                shutil.copyfileobj(fi, fo)
            o = oracle.Oracle.from_csv(csv)
            shutil.rmtree(d)
            r = 10.0
        else:
            import datagen
            labels, x, y, S, p = datagen.generate(name)
            fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
            o = oracle.Oracle.from_arrays(x.astype(np.float32), y.astype(np.float32), fc,
                                          np.array([0, 0, S, S], np.float32))
            del labels, x, y
            r = p["r"]
        o.bin(r)
        out = {"workload": name, "r": r, **o.boundary_pairs(), "seconds": round(time.time() - t0, 1)}
        print(json.dumps(out), flush=True)
        del o


if __name__ == "__main__":
    main()
