"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/native/oracle_sanitized_run.cpp drives every oracle entry point (each reference
function, whole one- and two-stage steps, diagnostics, single-point kernels) on small boxes
over three boundary sets and both energy fluxes.  It is compiled with the oracle's own
source under g++ -fsanitize=address,undefined and must run clean: an out-of-bounds index, a
use after free or a signed overflow in the checker would fail here, on the CPU.  A
ThreadSanitizer build is not run: the OpenMP runtime is not instrumented, and its
barriers read as races.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_runs_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_sanitized_run")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fopenmp", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-I", os.path.join(ROOT, "oracle"),
           os.path.join(HERE, "native", "oracle_sanitized_run.cpp"), os.path.join(ROOT, "oracle", "mhd_oracle.cpp"),
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "asan" in (r.stderr or "").lower():
        pytest.skip("the sanitizer runtime is not available to g++ here")
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    run = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    out = run.stdout + run.stderr
    assert run.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    assert "oracle sanitized run ok" in out
