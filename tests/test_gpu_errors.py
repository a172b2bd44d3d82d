"""Error behaviour of the C ABI on a GPU: bad arguments and calls out of order return a
negative status with a message (raised as MhdError by the Python layer), leave the context
usable, and a good call afterwards works.  The reference aborts on such input; the C++
drop-in layer turns these statuses into the same fail-fast abort."""
import numpy as np
import pytest

from mhdturbulence_amd import Batch, Box, Config, capi
from mhdturbulence_amd.capi import MhdError

pytestmark = pytest.mark.gpu

BASE = dict(bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5), gam=5.0 / 3.0, energy_flux=1,
            grid_shift=0)


@pytest.mark.parametrize("bad", [dict(nx=3), dict(bc=(1, 1, 2, 2, 1, 4)), dict(gam=1.0), dict(cfl=0.0),
                                 dict(energy_flux=2), dict(precision=3), dict(stages=3),
                                 dict(xmax=(0.5, -0.5, 0.5))])
def test_create_rejects_bad_config(bad):
    cfg = dict(nx=16, ny=16, nz=16, **BASE)
    cfg.update(bad)
    with pytest.raises(MhdError) as e:
        Box(Config(**cfg))
    assert "(-1)" in str(e.value)  # MHD_ERR_ARG


def test_calls_out_of_order_and_bad_arguments():
    b = Box(Config(nx=16, ny=16, nz=16, **BASE))
    with pytest.raises(MhdError) as e:  # no grid yet
        b.run(2)
    assert "(-5)" in str(e.value) and "grid" in str(e.value)
    with pytest.raises(MhdError):
        b.control_timestep()
    b.generate_grid()
    b.generate_problem_shear()
    with pytest.raises(MhdError):
        b.run(2, mode=2)
    with pytest.raises(MhdError):
        b.run(-1)
    with pytest.raises(MhdError):
        b.get_numerical_flux(4, 1.0)
    with pytest.raises(MhdError):  # fluxes of a staged step were never computed
        b.download_flux(1)
    # the context still works
    t, dts, done = b.run(3)
    assert done == 3 and t > 0 and np.all(dts > 0) and b.status() == 0
    b.close()


def test_batch_errors():
    bt = Batch(Config(**BASE), 2, 1 << 28)
    with pytest.raises(MhdError):
        bt.add_box(2, 16, 16)
    with pytest.raises(MhdError):
        bt.add_box(16, 16, 16, ic=dict(a=0.0))
    ids = [bt.add_box(16, 16, 16, ic={}), bt.add_box(16, 16, 16, ic={})]
    with pytest.raises(MhdError):  # no free slot
        bt.add_box(16, 16, 16, ic={})
    assert bt.try_add_box(16, 16, 16, ic={}) is None
    with pytest.raises(MhdError):
        bt.remove_box(7)
    bt.remove_box(ids[0])
    bt.add_box(16, 16, 16, ic={})
    bt.step(2)
    bt.close()


def test_error_text_api():
    L = capi.load()
    assert L.mhd_error_string(-1).decode()
    assert L.mhd_version() > 0
