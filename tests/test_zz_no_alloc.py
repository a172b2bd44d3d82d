"""mhd_run_noalloc (the allocation-free fused run) against mhd_run, through the C ABI: two
calls in a row, the second on a smaller box, each bitwise equal to mhd_run on the same
state; the context's device bytes do not grow; boxes above the static scratch are refused
with a message; runs longer than one dt-log chunk continue as mhd_run does."""
import ctypes as C
import sys

import numpy as np
import pytest

from mhdturbulence_amd import Box, Config
from mhdturbulence_amd.capi import MhdError

pytestmark = pytest.mark.gpu

BASE = dict(bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5), gam=5.0 / 3.0, energy_flux=1,
            grid_shift=0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _box(shape, **kw):
    b = Box(Config(nx=shape[0], ny=shape[1], nz=shape[2], **BASE, **kw))
    b.generate_grid()
    b.generate_problem_shear(bx0=0.1, amp=0.02, a=0.06)
    return b


def _noalloc(b, nsteps, time=0.0):
    """the C-ABI call itself"""
    t = C.c_double(time)
    log = np.zeros(max(nsteps, 1))
    done = C.c_int()
    rc = b.L.mhd_run_noalloc(b.h, nsteps, 0.0, C.byref(t), log.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done))
    assert rc == 0, b.L.mhd_last_error()
    return t.value, log[:done.value].copy(), done.value


@pytest.mark.parametrize("kw", [dict(), dict(precision=2), dict(stages=2, cfl=0.3)])
def test_noalloc_equals_run_twice_second_smaller(kw):
    for shape, steps in (((48, 40, 32), 21), ((24, 20, 16), 13)):
        say("noalloc vs run", shape, kw)
        ref = _box(shape, **kw)
        t_r, log_r, done_r = ref.run(steps, mode=0)
        b = _box(shape, **kw)
        bytes0 = b.device_bytes()
        t_n, log_n, done_n = _noalloc(b, steps)
        assert b.device_bytes() == bytes0  # no lazily allocated buffers
        assert done_n == done_r == steps and t_n == t_r
        assert np.array_equal(log_n, log_r)
        assert np.array_equal(b.download_cons(), ref.download_cons())
        assert np.array_equal(b.download_prim(), ref.download_prim())
        ref.close()
        b.close()


def test_noalloc_longer_than_one_log_chunk():
    shape, steps = (16, 16, 8), 4100
    say("noalloc", steps, "steps")
    ref = _box(shape)
    t_r, log_r, _ = ref.run(steps)
    b = _box(shape)
    t_n, log_n, done = _noalloc(b, steps)
    assert done == steps
    np.testing.assert_allclose(log_n, log_r, rtol=1e-13)
    assert t_n == pytest.approx(t_r, rel=1e-13)
    np.testing.assert_allclose(b.interior(b.download_cons()), ref.interior(ref.download_cons()), rtol=1e-11,
                               atol=1e-13)


def test_noalloc_refuses_box_above_scratch():
    b = _box((96, 96, 96))
    with pytest.raises(MhdError) as e:
        b.run_noalloc(2)
    assert "(-1)" in str(e.value) and "MHD_NOALLOC_SCRATCH_BYTES" in str(e.value)
    # the context is untouched and still runs through mhd_run
    assert b.run(2)[2] == 2
    b.close()
