"""Asynchronous snapshots (mhd_snapshot_begin / _end): the primitive state copied to pinned
host memory on a second stream while the following steps run equals a synchronous download
taken at the same step, and the run itself is not disturbed."""
import numpy as np
import pytest

from mhdturbulence_amd import Box, Config
from mhdturbulence_amd.capi import MhdError

pytestmark = pytest.mark.gpu

BASE = dict(bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5), gam=5.0 / 3.0, energy_flux=1,
            grid_shift=0)


def _box(precision=0):
    b = Box(Config(nx=48, ny=40, nz=32, precision=precision, **BASE))
    b.generate_grid()
    b.generate_problem_shear(bx0=0.1, a=0.05, amp=0.02, seed=5)
    return b


@pytest.mark.parametrize("precision", [0, 1])
def test_snapshot_overlaps_steps_and_equals_download(precision):
    a, r = _box(precision), _box(precision)
    t, _, _ = a.run(5)
    a.snapshot_begin()
    t2, _, _ = a.run(7, time=t)            # runs while the snapshot transfers
    Ps = a.snapshot_end()
    tr, _, _ = r.run(5)
    assert tr == t
    np.testing.assert_array_equal(Ps, r.download_prim())
    tr2, _, _ = r.run(7, time=tr)
    assert tr2 == t2
    np.testing.assert_array_equal(a.download_prim(), r.download_prim())
    # a second begin while the first is in flight waits for it; zero-copy view of the latest
    a.snapshot_begin()
    t3, _, _ = a.run(3, time=t2)
    a.snapshot_begin()
    view = a.snapshot_end(copy=False)
    r.run(3, time=tr2)
    np.testing.assert_array_equal(view, r.download_prim())
    a.close()
    r.close()


def test_snapshot_end_without_begin_is_refused():
    b = _box()
    with pytest.raises(MhdError):
        b.snapshot_end()
    b.snapshot_begin()
    b.snapshot_end()
    with pytest.raises(MhdError):
        b.snapshot_end()
    b.close()
