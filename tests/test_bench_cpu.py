"""bench.py's contract checks that need no GPU: more than one GPU is refused with a message
before any device work, and the option parser knows every documented flag."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout


def test_documented_flags_are_accepted():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--n", "--shape", "--mode", "--precision", "--no-cpu", "--no-serve",
                 "--profile-steps", "--cpu-steps", "--serve-requests"):
        assert flag in r.stdout, flag


def test_cpu_baseline_binds_threads_and_names_the_cpu():
    """the baseline leg runs the oracle in a child process with one thread per physical core
    (at most OMP_NUM_THREADS), close binding on cores, and reports the CPU it ran on"""
    sys.path.insert(0, ROOT)
    import bench
    line = bench.cpu_baseline((16, 12, 8), 2)
    cpu = bench.host_cpu()
    assert line["kind"] == "port" and line["unit"] == "cell-updates/s" and line["value"] > 0
    assert 1 <= line["cores"] <= cpu["physical_cores"] <= cpu["logical_cpus"]
    assert line["omp"] == {"OMP_NUM_THREADS": line["cores"], "OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}
    assert line["cpu"]["model"] == cpu["model"]
