"""Parity at the workloads the benchmark and the configurations are quoted on.

* The benchmark's own box (bench.make_box): 256^3 KH shear layer, gamma 5/3, Bx 0.1,
  vy kick from seed 20261203, energy flux (v.B)B_n, CFL factor 0.05.  At 256^3 the fused
  kernels take tilings no smaller test reaches (FP64: 43-row y chunks of four-strip blocks
  with the shared edge pre-pass and 64-plane z chunks; mixed: 43-row y chunks of
  block-cooperative FP32 strip edges and 64-plane z chunks).  The GPU's device-generated initial state is copied
  into the oracle, so both start from the same bits; then
  - FP64, fused kernels, 10 steps: rel L1 <= 1e-10, dt sequences equal to 1e-12;
  - the two-stage Runge-Kutta step at CFL 0.3, 5 steps: the same bar;
  - precision 1 and 2, 10 steps: rel L1 of rho and p <= 1e-6, every variable <= 1e-4.
* Config 2's physics at 150^3 (the same shear layer, 150^3 cells), fused and staged paths,
  100 steps against the oracle at the FP64 bar.

This is the loop body of src_cpp_kokkos_host/main.cpp:229-259 at the metric's shape.
"""
import sys

import numpy as np
import pytest

import bench
import parity

pytestmark = pytest.mark.gpu

TOL = 1e-10
SHAPE = (256, 256, 256)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_like(orc, shape, stages=1, cfl=0.05):
    """an oracle box with bench.make_box's configuration"""
    o = orc.Oracle(nx=shape[0], ny=shape[1], nz=shape[2], bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -0.5, -0.5),
                   xmax=(0.5, 0.5, 0.5), gam=5.0 / 3.0, cfl=cfl, energy_flux=1, grid_shift=0, stages=stages)
    o.generate_grid()
    return o


@pytest.fixture(scope="module")
def initial():
    """the benchmark box's initial state as the device generates it (P, U in the host layout)"""
    say("headline: generating the 256^3 benchmark box on the device")
    b = bench.make_box(SHAPE, 0)
    P, U = b.download_prim(), b.download_cons()
    b.close()
    return P, U


@pytest.fixture(scope="module")
def oracle_10(oracle_lib, initial):
    say("headline: oracle, 10 steps of 256^3")
    o = oracle_like(oracle_lib, SHAPE)
    o.P[...] = initial[0]
    o.U[...] = initial[1]
    t, log = o.run(10)
    return o.interior(o.P), t, log.copy()


def _gpu_run(initial, nsteps, precision=0, stages=1, cfl=0.05):
    b = bench.make_box(SHAPE, precision, stages, cfl)
    b.upload_prim(initial[0])
    b.upload_cons(initial[1])
    t, log, done = b.run(nsteps, mode=0)
    assert done == nsteps and b.status() == 0
    P = b.interior(b.download_prim())
    b.close()
    return P, t, log


def test_headline_256cube_fp64_fused_10_steps(initial, oracle_10):
    say("headline: FP64 fused, 10 steps")
    Po, to, logo = oracle_10
    Pg, tg, logg = _gpu_run(initial, 10)
    np.testing.assert_allclose(logg, logo, rtol=1e-12)
    assert tg == pytest.approx(to, rel=1e-12)
    err = parity.rel_l1(Pg, Po, "prim")
    say("headline FP64 rel L1", err.max())
    assert err.max() <= TOL, err


@pytest.mark.parametrize("precision", [1, 2])
def test_headline_256cube_mixed_10_steps(initial, oracle_10, precision):
    say("headline: precision", precision)
    Po, _, logo = oracle_10
    Pg, _, logg = _gpu_run(initial, 10, precision=precision)
    err = parity.rel_l1(Pg, Po, "prim")
    say("headline mixed rel L1", precision, err)
    assert err[0] <= 1e-6 and err[5] <= 1e-6, err
    assert np.all(err <= 1e-4), err
    np.testing.assert_allclose(logg, logo, rtol=1e-6)


def test_headline_256cube_two_stage_cfl03_5_steps(oracle_lib, initial):
    say("headline: two-stage step at CFL 0.3, oracle 5 steps")
    o = oracle_like(oracle_lib, SHAPE, stages=2, cfl=0.3)
    o.P[...] = initial[0]
    o.U[...] = initial[1]
    to, logo = o.run(5)
    Po = o.interior(o.P)
    del o
    Pg, tg, logg = _gpu_run(initial, 5, stages=2, cfl=0.3)
    np.testing.assert_allclose(logg, logo, rtol=1e-12)
    err = parity.rel_l1(Pg, Po, "prim")
    say("headline RK2 rel L1", err.max())
    assert err.max() <= TOL, err


def test_config2_shear_150cube_100_steps(oracle_lib):
    """config 2's physics at 150^3 (32-cell strips, the last one ragged), both paths"""
    shape = (150, 150, 150)
    say("config 2 shear layer 150^3: oracle 100 steps")
    b = bench.make_box(shape, 0)
    P0, U0 = b.download_prim(), b.download_cons()
    o = oracle_like(oracle_lib, shape)
    o.P[...] = P0
    o.U[...] = U0
    to, logo = o.run(100)
    Po = o.interior(o.P)
    for mode in (0, 1):
        say("config 2 150^3 mode", mode)
        b.upload_prim(P0)
        b.upload_cons(U0)
        tg, logg, done = b.run(100, mode=mode)
        assert done == 100 and b.status() == 0
        np.testing.assert_allclose(logg, logo, rtol=1e-12)
        err = parity.rel_l1(b.interior(b.download_prim()), Po, "prim")
        say("150^3 rel L1", mode, err.max())
        assert err.max() <= TOL, (mode, err)
    b.close()
