"""The rocprofv3 summary tool on synthetic CSV files (CPU only).

- `kernels` sums a kernel's dispatches across template instantiations.
- `traffic` scales FETCH_SIZE / WRITE_SIZE by the plain copy's reading.
- `counters` joins a trace run and counter runs whose dispatch ids differ,
  on the n-th dispatch of each kernel.
"""
import csv
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TOOL = os.path.join(HERE, "..", "gpu-colocationmininig_amd", "profile_summary.py")


def _write(path, fields, rows):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)
    return str(path)


def _trace(path, disp):
    """disp: (dispatch id, kernel name, start ns, end ns)."""
    return _write(path, ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"],
                  [dict(Dispatch_Id=d, Kernel_Name=n, Start_Timestamp=s, End_Timestamp=e) for d, n, s, e in disp])


def _pmc(path, rows):
    """rows: (dispatch id, kernel name, counter, value)."""
    return _write(path, ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"],
                  [dict(Dispatch_Id=d, Kernel_Name=n, Counter_Name=c, Counter_Value=v) for d, n, c, v in rows])


def _run(*args):
    r = subprocess.run([sys.executable, TOOL] + [str(a) for a in args], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_kernels_sums_instantiations(tmp_path):
    t = _trace(tmp_path / "t.csv", [(1, "void k_a<1>(int)", 0, 2_000_000), (2, "void k_a<2>(int)", 0, 1_000_000),
                                    (3, "k_b(int)", 0, 500_000)])
    out = _run("kernels", t)
    line = [l for l in out.splitlines() if "`k_a`" in l][0]
    assert "| 2 | 3000.0 | 1500.0 |" in line
    per = _run("kernels", t, "--templates")
    assert "`k_a<1>`" in per and "`k_a<2>`" in per


def test_traffic_scales_by_copy(tmp_path):
    copy = 256 * 1024 * 1024
    f = _pmc(tmp_path / "f.csv", [(5, "k_x(int)", "FETCH_SIZE", 1000.0), (9, "k_x(int)", "FETCH_SIZE", 3000.0)])
    w = _pmc(tmp_path / "w.csv", [(5, "k_x(int)", "WRITE_SIZE", 10.0), (9, "k_x(int)", "WRITE_SIZE", 30.0)])
    cf = _pmc(tmp_path / "cf.csv", [(1, "__amd_rocclr_copyBuffer", "FETCH_SIZE", copy / 1024 / 2)])
    cw = _pmc(tmp_path / "cw.csv", [(1, "__amd_rocclr_copyBuffer", "WRITE_SIZE", copy / 1024)])
    d = json.loads(_run("traffic", "--fetch", f, "--write", w, "--cfetch", cf, "--cwrite", cw, "--kernel", "k_x"))
    assert d["dispatches"] == 2
    assert d["fetch_scale"] == 2.0 and d["write_scale"] == 1.0
    assert d["hbm_bytes_per_launch"] == 2000.0 * 1024 * 2 + 20.0 * 1024


def test_counters_align_on_kernel_ordinal(tmp_path):
    # the trace run has an extra copy dispatch, so every later id is shifted
    t = _trace(tmp_path / "t.csv", [(1, "k_count(int)", 0, 3_000_000), (2, "copy", 0, 1000),
                                    (3, "k_count(int)", 0, 2_000_000)])
    p = _pmc(tmp_path / "p.csv", [(10, "k_count(int)", "GRBM_GUI_ACTIVE", 8000.0),
                                  (10, "k_count(int)", "SQ_INSTS_VALU", 111.0),
                                  (11, "k_count(int)", "GRBM_GUI_ACTIVE", 8000.0),
                                  (11, "k_count(int)", "SQ_INSTS_VALU", 222.0)])
    f = _pmc(tmp_path / "f.csv", [(20, "k_count(int)", "FETCH_SIZE", 1.0), (21, "k_count(int)", "FETCH_SIZE", 2.0)])
    w = _pmc(tmp_path / "w.csv", [(30, "k_count(int)", "WRITE_SIZE", 1.0), (31, "k_count(int)", "WRITE_SIZE", 2.0)])
    cf = _pmc(tmp_path / "cf.csv", [(1, "__amd_rocclr_copyBuffer", "FETCH_SIZE", 1024.0)])
    cw = _pmc(tmp_path / "cw.csv", [(1, "__amd_rocclr_copyBuffer", "WRITE_SIZE", 1024.0)])
    out = _run("counters", "--trace", t, "--pmc", p, "--fetch", f, "--write", w, "--cfetch", cf, "--cwrite", cw,
               "--match", "k_count", "--min-ms", "1", "--copy-bytes", 1024 * 1024)
    rows = [l for l in out.splitlines() if l.startswith("| `k_count`")]
    assert len(rows) == 2
    assert "| 3.0 |" in rows[0] and "| 111 |" in rows[0]
    assert "| 2.0 |" in rows[1] and "| 222 |" in rows[1]
