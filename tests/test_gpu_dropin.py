"""The C++ drop-in layer (mhdturbulence_amd/host: hydflux_mod / boundary_mod over the C ABI)
driving the src_cpp_kokkos_host call sequence on the 150^3 Kelvin-Helmholtz problem,
compared with the oracle after a few steps."""
import os
import subprocess

import numpy as np
import pytest

import parity

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mhdturbulence_amd", "mhd_kh_main")
N = 150
STEPS = 5


@pytest.fixture(scope="module")
def oracle_state(oracle_lib):
    o = oracle_lib.Oracle(nx=N, ny=N, nz=N)
    o.generate_grid()
    o.generate_problem_kh(vy_amp=1e-3, rrv=1e-2)
    t, log = o.run(STEPS)
    return o, t, log


@pytest.mark.parametrize("flag", [[], ["--staged"], ["--fused"], ["--peek", "2"]])
def test_kh_main_matches_oracle(tmp_path, oracle_state, flag):
    """the reference loop through the drop-in: recorded and run as the fused kernels (no
    flag), every function on its staged kernels (--staged), AdvanceSteps (--fused), and a
    recorded step interrupted by a flux download that runs out staged (--peek)"""
    o, t, log = oracle_state
    dump = tmp_path / "P.bin"
    r = subprocess.run([EXE, "--steps", str(STEPS), "--dump", str(dump), *flag], cwd=tmp_path, capture_output=True,
                       text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    P = np.fromfile(dump, dtype=np.float64).reshape(o.ktot, o.jtot, 12, o.itot)
    err = parity.rel_l1(o.interior(P), o.interior(o.P), "prim")
    assert np.all(err <= 1e-10), err
    # dt.log carries the first dt in %.17g
    first = open(tmp_path / "dt.log").readline()
    dt0 = float(first.split("dt=")[1])
    assert dt0 == pytest.approx(log[0], rel=1e-14)


def test_kh_main_256_matches_oracle(tmp_path, oracle_lib):
    """the same loop at 256^3 (mhd_kh_main_256, the reference KH problem with its grid size
    compiled in), recorded and run as the fused kernels: at this shape the FP64 xy kernel is
    the two-wavefront march with block-shared strip edges"""
    exe = os.path.join(ROOT, "mhdturbulence_amd", "mhd_kh_main_256")
    steps = 3
    o = oracle_lib.Oracle(nx=256, ny=256, nz=256)
    o.generate_grid()
    o.generate_problem_kh(vy_amp=1e-3, rrv=1e-2)
    t, log = o.run(steps)
    dump = tmp_path / "P256.bin"
    r = subprocess.run([exe, "--steps", str(steps), "--dump", str(dump)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    P = np.fromfile(dump, dtype=np.float64).reshape(o.ktot, o.jtot, 12, o.itot)
    err = parity.rel_l1(o.interior(P), o.interior(o.P), "prim")
    assert np.all(err <= 1e-10), err
    first = open(tmp_path / "dt.log").readline()
    assert float(first.split("dt=")[1]) == pytest.approx(log[0], rel=1e-14)
