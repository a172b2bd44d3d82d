"""The two-stage Runge-Kutta step (stages = 2) on the GPU against the oracle's.

The reference steps forward Euler only; stages = 2 wraps that stage as U* = E(U^n),
U^{n+1} = (U^n + E(U*)) / 2 with the dt and c_h of U^n.  The oracle's two-stage step is
pinned on the CPU (second-order convergence at fixed CFL, the psi-damping closed form);
here the fused and staged GPU paths, single boxes and batches, are held to the same bar as
the Euler step: per-variable relative L1 <= 1e-10 and equal dt sequences.
"""
import sys

import numpy as np
import pytest

import parity
from mhdturbulence_amd import Batch, Box, Config

pytestmark = pytest.mark.gpu

BASE = dict(bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5), gam=5.0 / 3.0, energy_flux=1,
            grid_shift=0, cfl=0.3, stages=2)
IC = dict(a=0.06, bx0=0.12, amp=0.03)
TOL_L1 = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def orc(oracle_lib):
    return oracle_lib


def _pair(orc, n, ic=IC, **kw):
    cfg = {**BASE, **kw}
    o = orc.Oracle(nx=n[0], ny=n[1], nz=n[2], **{k: v for k, v in cfg.items() if k != "precision"})
    o.generate_grid()
    o.generate_problem_shear(**ic)
    b = Box(Config(nx=n[0], ny=n[1], nz=n[2], **cfg))
    b.set_grid(o.grid(1), o.grid(1, True), o.grid(2), o.grid(2, True), o.grid(3), o.grid(3, True))
    b.upload_prim(o.P)
    b.upload_cons(o.U)
    return o, b


def _errs(o, b):
    errP = parity.rel_l1(o.interior(b.download_prim()), o.interior(o.P), "prim")
    errU = parity.rel_l1(o.interior(b.download_cons()), o.interior(o.U), "cons")
    return errP, errU


@pytest.mark.parametrize("mode", [0, 1])
def test_rk2_100_steps_match_oracle(orc, mode):
    say("rk2 32^3 100 steps mode", mode)
    o, b = _pair(orc, (32, 32, 32))
    t_o, log_o = o.run(100)
    t_g, log_g, done = b.run(100, mode=mode)
    assert done == 100
    np.testing.assert_allclose(log_g, log_o, rtol=1e-12)
    assert t_g == pytest.approx(t_o, rel=1e-12)
    errP, errU = _errs(o, b)
    say(f"  max relL1 P={errP.max():.3e} U={errU.max():.3e}")
    assert np.all(errP <= TOL_L1) and np.all(errU <= TOL_L1), (errP, errU)


@pytest.mark.parametrize("shape,tiling", [((24, 20, 16), (16, 4, 3)), ((17, 9, 7), None), ((65, 6, 5), None)])
def test_rk2_chunks_and_ragged_boxes(orc, shape, tiling):
    """several y and z chunks per column, narrow and ragged strips: the second stage reads
    the first stage's output and its own cell's U^n across every chunk seam"""
    say("rk2 chunks", shape, tiling)
    for mode in (0, 1):
        o, b = _pair(orc, shape, energy_flux=0)
        if tiling and mode == 0:
            b.set_tiling(*tiling)
        o.run(20)
        _, log, done = b.run(20, mode=mode)
        assert done == 20
        errP, errU = _errs(o, b)
        assert np.all(errP <= TOL_L1) and np.all(errU <= TOL_L1), (mode, errP, errU)


@pytest.mark.parametrize("bc", [(3, 3, 2, 2, 3, 3), (2, 2, 3, 3, 1, 1)])
def test_rk2_other_boundaries(orc, bc):
    """outflow and reflecting walls: the ghost images of each stage's output (stage 2: of the
    averaged state) are what the next stage and step read"""
    say("rk2 boundaries", bc)
    for mode in (0, 1):
        o, b = _pair(orc, (24, 24, 20), bc=bc)
        o.run(25)
        _, _, done = b.run(25, mode=mode)
        assert done == 25
        errP, errU = _errs(o, b)
        assert np.all(errP <= TOL_L1) and np.all(errU <= TOL_L1), (mode, errP, errU)


@pytest.mark.parametrize("shape,bc", [((256, 12, 8), (1, 1, 2, 2, 1, 1)), ((128, 10, 6), (1, 1, 3, 3, 3, 3)),
                                      ((192, 9, 4), (1, 1, 2, 2, 2, 2))])
def test_rk2_two_waves_per_simd(orc, shape, bc):
    """both stages on the FP64 xy march at two wavefronts per SIMD (block-shared strip edges),
    forced on small boxes, with reflecting and outflow walls in y and z"""
    say("rk2 two wavefronts per SIMD", shape, bc)
    o, b = _pair(orc, shape, bc=bc)
    b.set_xy_waves_per_simd(2)
    assert b.get_xy_waves_per_simd() == 2
    t_o, log_o = o.run(20)
    t_g, log_g, done = b.run(20, mode=0)
    assert done == 20
    np.testing.assert_allclose(log_g, log_o, rtol=1e-12)
    errP, errU = _errs(o, b)
    say("  relL1", errP.max(), errU.max())
    assert errP.max() <= TOL_L1 and errU.max() <= TOL_L1


def test_rk2_run_continues_across_calls_and_stops_at_tmax(orc):
    o, b = _pair(orc, (24, 24, 16))
    t1, _, d1 = b.run(7, mode=0)
    t2, _, d2 = b.run(9, time=t1, mode=0)
    t_o, log_o = o.run(16)
    assert d1 == 7 and d2 == 9 and t2 == pytest.approx(t_o, rel=1e-12)
    errP, errU = _errs(o, b)
    assert np.all(errP <= TOL_L1), errP
    # tmax: the step that passes tmax is taken, then the loop stops
    o2, b2 = _pair(orc, (24, 24, 16))
    _, log2 = o2.run(12)
    tmax = float(np.sum(log2[:5])) + 0.5 * log2[5]
    t, _, done = b2.run(50, tmax=tmax, mode=0)
    assert done == 6 and t == pytest.approx(float(np.sum(log2[:6])), rel=1e-12)


def test_rk2_mixed_precision_close_to_fp64(orc):
    o, b = _pair(orc, (32, 32, 32), precision=1)
    o.run(60)
    b.run(60, mode=0)
    errP, _ = _errs(o, b)
    say(f"rk2 mixed vs fp64 oracle: max relL1 P={errP.max():.3e}")
    assert errP[0] < 1e-6 and errP[5] < 1e-6 and errP.max() < 1e-4, errP


def test_rk2_profile_names_fold_both_stages(orc):
    _, b = _pair(orc, (32, 16, 16))
    prof = b.profile_kernels(3, mode=0)
    assert list(prof) == ["fused_xy", "fused_z_update", "step_end"] and all(v > 0 for v in prof.values())
    prof = b.profile_kernels(3, mode=1)
    assert len(prof) == 8 and all(v > 0 for v in prof.values())


def test_rk2_batch_matches_oracle_and_single_box(orc):
    shapes = [(32, 32, 32), (48, 32, 16), (24, 40, 32)]
    ics = [dict(a=0.05, bx0=0.1, amp=0.01, p0=1.0, seed=11), dict(a=0.08, bx0=0.0, amp=0.005, p0=2.5, seed=12),
           dict(a=0.03, bx0=0.2, amp=0.002, p0=0.7, seed=13)]
    bt = Batch(Config(**BASE), 4, 2 << 30)
    ids = [bt.add_box(*s, ic=ic) for s, ic in zip(shapes, ics)]
    bt.step(15)
    for bid, s, ic in zip(ids, shapes, ics):
        o = orc.Oracle(nx=s[0], ny=s[1], nz=s[2], **BASE)
        o.generate_grid()
        o.generate_problem_shear(**ic)
        t, _ = o.run(15)
        st = bt.box_state(bid)
        assert st["steps"] == 15 and st["time"] == pytest.approx(t, rel=1e-12)
        _, P = bt.finish_box(bid, want_prim=True)
        err = parity.rel_l1(o.interior(P), o.interior(o.P), "prim")
        say("rk2 batch", s, "max relL1", err.max())
        assert err.max() < TOL_L1, (s, err)
    # the batch's box equals the single-box path on the device-generated state
    s, ic = shapes[1], ics[1]
    x = Box(Config(nx=s[0], ny=s[1], nz=s[2], **BASE))
    x.generate_grid()
    x.generate_problem_shear(**ic)
    x.run(15, mode=0)
    bt2 = Batch(Config(**BASE), 2, 1 << 30)
    bid = bt2.add_box(*s, ic=ic)
    bt2.step(15)
    _, Pb = bt2.finish_box(bid, want_prim=True)
    assert parity.rel_l1(x.interior(Pb), x.interior(x.download_prim()), "prim").max() < 1e-13


def test_rk2_mixed_batch_equals_single_box():
    """mixed precision, two stages: a box in a batch (beside boxes of other sizes) matches the
    same box stepped alone by the single-box path.  The two run different chunk lengths, and
    FP32 faces computed in different code paths may contract differently, so the bound is
    FP32 round-off over the steps rather than bitwise"""
    s, ic = (48, 32, 16), dict(a=0.08, bx0=0.1, amp=0.005, p0=2.5, seed=12)
    cfg = Config(**BASE, precision=1)
    x = Box(Config(nx=s[0], ny=s[1], nz=s[2], **BASE, precision=1))
    x.generate_grid()
    x.generate_problem_shear(**ic)
    x.run(12, mode=0)
    bt = Batch(cfg, 4, 2 << 30)
    bt.add_box(32, 32, 32, ic=dict(a=0.05, seed=11))
    bid = bt.add_box(*s, ic=ic)
    bt.add_box(24, 40, 32, ic=dict(a=0.03, seed=13))
    bt.step(12)
    _, Pb = bt.finish_box(bid, want_prim=True)
    err = parity.rel_l1(x.interior(Pb), x.interior(x.download_prim()), "prim").max()
    say(f"rk2 mixed batch vs single box: max relL1 {err:.3e}")
    assert err < 1e-8


def test_rk2_fp32_divergence_close_to_fp64(orc):
    o, b = _pair(orc, (32, 32, 32), precision=2)
    o.run(40)
    b.run(40, mode=0)
    errP, _ = _errs(o, b)
    say(f"rk2 precision 2 vs fp64 oracle: max relL1 P={errP.max():.3e}")
    assert errP[0] < 1e-6 and errP[5] < 1e-6 and errP.max() < 1e-4, errP
