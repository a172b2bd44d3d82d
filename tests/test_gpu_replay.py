"""Deterministic replay and checkpoint / resume on the GPU.

The reductions of the step are order-independent (integer atomicMin / atomicMax on the bit
patterns of non-negative doubles for dt and c_h; the t-prof sums over a fixed tree), and no
kernel reads what another block of the same launch writes.  So a run repeated from the same
state gives the same bits, whatever order the blocks ran in.  The reference has no restart
reader; downloading P and U (the host arrays of src_cpp_kokkos_host/mhd.hpp:269-275) and
uploading them into a fresh box is this library's checkpoint, and the resumed run must be
the uninterrupted one bit for bit.
"""
import sys

import numpy as np
import pytest

import bench
from mhdturbulence_amd import Batch, Box, Config

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(shape, steps, precision=0, stages=1):
    b = bench.make_box(shape, precision, stages, 0.3 if stages == 2 else 0.05)
    t, log, done = b.run(steps, mode=0)
    assert done == steps and b.status() == 0
    P = b.download_prim()
    d = b.diagnostics()
    b.close()
    return t, log, P, d


@pytest.mark.parametrize("shape,precision,stages", [((128, 128, 128), 0, 1), ((128, 96, 64), 2, 1),
                                                     ((96, 80, 48), 0, 2)])
def test_repeated_run_is_bitwise_identical(shape, precision, stages):
    """two runs of the same device-generated state: equal dt logs, primitives and t-prof
    diagnostics, bit for bit (128^3 FP64 takes the two-wavefront xy march)"""
    say("replay", shape, "precision", precision, "stages", stages)
    a = _run(shape, 30, precision, stages)
    b = _run(shape, 30, precision, stages)
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])
    assert np.array_equal(a[3], b[3])


def test_batch_replay_is_bitwise_identical():
    """a batch of uneven boxes stepped twice from the same states: every box's primitives
    are the same bits"""
    from mhdturbulence_amd.serve import ensemble_box

    def once():
        cfg = Config(**bench.SHEAR_BASE, precision=0)
        bt = Batch(cfg, max_boxes=6, arena_bytes=2 << 30)
        ids = []
        for k in range(6):
            (nx, ny, nz), ic = ensemble_box(k)
            ids.append(bt.add_box(nx // 2, ny // 2, nz // 2, ic=ic))
        bt.step(15)
        out = [bt.finish_box(i, want_prim=True)[1] for i in ids]
        bt.close()
        return out
    say("batch replay")
    a, b = once(), once()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("shape", [(128, 128, 128), (48, 40, 32)])
def test_checkpoint_resume_is_bitwise_identical(shape):
    """20 steps, download P and U and the time, upload them into a fresh box, 10 more steps:
    the same bits and dt sequence as 30 uninterrupted steps"""
    say("checkpoint / resume", shape)
    ref = bench.make_box(shape, 0)
    t_ref, log_ref, done = ref.run(30, mode=0)
    assert done == 30
    P_ref = ref.download_prim()
    ref.close()

    a = bench.make_box(shape, 0)
    t20, log20, _ = a.run(20, mode=0)
    P, U = a.download_prim(), a.download_cons()
    a.close()
    b = bench.make_box(shape, 0)
    b.upload_prim(P)
    b.upload_cons(U)
    t30, log10, done = b.run(10, time=t20, mode=0)
    assert done == 10
    np.testing.assert_array_equal(np.concatenate([log20, log10]), log_ref)
    assert t30 == t_ref
    np.testing.assert_array_equal(b.download_prim(), P_ref)
    b.close()
