"""Batched boxes (mhd_batch_*): per-box results equal the oracle and equal the same box run
alone, each box keeps its own dt sequence, and retiring / admitting boxes reuses the arena."""
import sys

import numpy as np
import pytest

import parity
from mhdturbulence_amd import Batch, Box, Config

pytestmark = pytest.mark.gpu

BASE = dict(bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5), gam=5.0 / 3.0, energy_flux=1,
            grid_shift=0)
SHAPES = [(32, 32, 32), (48, 32, 16), (24, 40, 32), (64, 16, 24)]
ICS = [dict(a=0.05, bx0=0.1, amp=0.01, p0=1.0, seed=11), dict(a=0.08, bx0=0.0, amp=0.005, p0=2.5, seed=12),
       dict(a=0.03, bx0=0.2, amp=0.002, p0=0.7, seed=13), dict(a=0.06, bx0=0.15, amp=0.008, p0=4.0, seed=14)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _batch(max_boxes=8, gib=2.0, precision=0):
    return Batch(Config(**BASE, precision=precision), max_boxes, int(gib * 2**30))


@pytest.fixture(scope="module")
def orc(oracle_lib):
    return oracle_lib


def test_batch_boxes_match_oracle_and_own_dt(orc):
    say("batch vs oracle")
    b = _batch()
    ids = [b.add_box(*s, ic=ic) for s, ic in zip(SHAPES, ICS)]
    b.step(25)
    for bid, s, ic in zip(ids, SHAPES, ICS):
        o = orc.Oracle(nx=s[0], ny=s[1], nz=s[2], **BASE)
        o.generate_grid()
        o.generate_problem_shear(**ic)
        t, log = o.run(25)
        st = b.box_state(bid)
        assert st["steps"] == 25 and not st["stopped"]
        assert st["time"] == pytest.approx(t, rel=1e-12)   # the box's own dt sequence
        _, P = b.finish_box(bid, want_prim=True)
        # device initial state vs libm initial state differ in the last bits only
        err = parity.rel_l1(o.interior(P), o.interior(o.P), "prim")
        say(s, "max relL1", err.max())
        assert err.max() < 1e-10, (s, err)


def test_batch_alone_equals_in_mix_bitwise():
    say("alone vs mix")
    alone = _batch()
    a = alone.add_box(*SHAPES[2], ic=ICS[2])
    alone.step(30)
    _, Pa = alone.finish_box(a, want_prim=True)
    mix = _batch()
    ids = [mix.add_box(*s, ic=ic) for s, ic in zip(SHAPES, ICS)]
    mix.step(30)
    _, Pm = mix.finish_box(ids[2], want_prim=True)
    np.testing.assert_array_equal(Pa, Pm)


def test_batch_equals_single_box_path():
    s, ic = SHAPES[1], ICS[1]
    b = _batch()
    bid = b.add_box(*s, ic=ic)
    b.step(20)
    _, Pb = b.finish_box(bid, want_prim=True)
    x = Box(Config(nx=s[0], ny=s[1], nz=s[2], **BASE))
    x.generate_grid()
    x.generate_problem_shear(**ic)
    x.run(20, mode=0)
    err = parity.rel_l1(x.interior(Pb), x.interior(x.download_prim()), "prim")
    assert err.max() < 1e-13, err


def test_batch_stop_retire_admit_reuses_arena():
    b = _batch(max_boxes=3, gib=0.5)
    need = b.box_bytes(32, 32, 32)
    largest0, total0 = b.free_bytes()
    ids = [b.add_box(32, 32, 32, tmax=0.002 * (k + 1), ic=dict(seed=k)) for k in range(3)]
    assert b.try_add_box(16, 16, 16) is None  # no free slot
    done = set()
    for _ in range(40):
        b.step(5)
        for bid in b.poll():
            if bid not in done:
                st = b.box_state(bid)
                assert st["stopped"] and st["time"] > 0.002 * (bid + 1)
                d, _ = b.finish_box(bid)
                assert np.isfinite(d).all()
                b.remove_box(bid)
                done.add(bid)
        if len(done) == 3:
            break
    assert done == set(ids)
    assert b.free_bytes() == (largest0, total0)  # every region returned and merged
    assert need * 3 <= total0
    nid = b.add_box(40, 24, 16)
    b.step(3)
    assert b.box_state(nid)["steps"] == 3


def test_batch_mixed_precision_close_to_fp64():
    bd, bm = _batch(), _batch(precision=1)
    i = bd.add_box(*SHAPES[0], ic=ICS[0])
    j = bm.add_box(*SHAPES[0], ic=ICS[0])
    bd.step(40)
    bm.step(40)
    _, Pd = bd.finish_box(i, want_prim=True)
    _, Pm = bm.finish_box(j, want_prim=True)
    s = (slice(2, -3), slice(2, -3), slice(None), slice(2, -3))
    err = parity.rel_l1(Pm[s].transpose(2, 0, 1, 3), Pd[s].transpose(2, 0, 1, 3), "prim")
    say("mixed vs fp64", err)
    assert err[[0, 5]].max() < 1e-6 and err.max() < 1e-4


def test_batch_mixed_shared_strip_edges_alone_equals_mix(orc):
    """mixed precision in a batch with boxes whose FP32 strips share their edge faces in the
    block (one, two and four 64-cell strips per x-row) beside a narrow-strip box: each box
    is bitwise the same alone and in the mix, and close to the FP64 oracle"""
    shapes = [(64, 16, 24), (128, 8, 10), (256, 6, 4), (48, 32, 16)]
    mix = _batch(precision=1)
    ids = [mix.add_box(*s, ic=ic) for s, ic in zip(shapes, ICS)]
    mix.step(20)
    for bid, s, ic in zip(ids, shapes, ICS):
        _, Pm = mix.finish_box(bid, want_prim=True)
        alone = _batch(precision=1)
        a = alone.add_box(*s, ic=ic)
        alone.step(20)
        _, Pa = alone.finish_box(a, want_prim=True)
        alone.close()
        np.testing.assert_array_equal(Pa, Pm)
        o = orc.Oracle(nx=s[0], ny=s[1], nz=s[2], **BASE)
        o.generate_grid()
        o.generate_problem_shear(**ic)
        o.run(20)
        err = parity.rel_l1(o.interior(Pm), o.interior(o.P), "prim")
        say(s, "mixed vs oracle", err[[0, 5]])
        assert err[[0, 5]].max() < 1e-6 and err.max() < 1e-4, (s, err)
    mix.close()


def test_batch_fp32_divergence_matches_single_box():
    """precision 2 in a batch (chunks of other lengths, boxes of other sizes beside it)
    against the same box run alone: FP32 round-off apart"""
    s, ic = SHAPES[1], ICS[1]
    b = _batch(precision=2)
    ids = [b.add_box(*sh, ic=c) for sh, c in zip(SHAPES, ICS)]
    b.step(15)
    _, Pb = b.finish_box(ids[1], want_prim=True)
    x = Box(Config(nx=s[0], ny=s[1], nz=s[2], **BASE, precision=2))
    x.generate_grid()
    x.generate_problem_shear(**ic)
    x.run(15, mode=0)
    err = parity.rel_l1(x.interior(Pb), x.interior(x.download_prim()), "prim")
    say("precision 2 batch vs alone", err.max())
    assert err.max() < 1e-8, err
