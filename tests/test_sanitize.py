"""CPU sanitizer runs: the oracle (oracle/colo_oracle.c) and the host half of
libcolo (gpu-colocationmininig_amd/csrc/colo_host.cpp: CSV parser, binary
cache, candidate generation), each built with AddressSanitizer and
UndefinedBehaviorSanitizer by tests/sanitize/Makefile and driven through
valid and malformed inputs.  A sanitizer report or a failed check fails the
test.  No GPU code is built."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="no host compiler")
@pytest.mark.parametrize("prog", ["oracle_san", "host_san"])
def test_sanitized_build_runs_clean(prog, tmp_path):
    subprocess.run(["make", "-s", "-C", HERE, os.path.join(HERE, "_build", prog)], check=True, capture_output=True,
                   timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([os.path.join(HERE, "_build", prog), str(tmp_path)], capture_output=True, text=True,
                       timeout=120, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out, out
    assert "sanitizer run ok" in r.stdout
