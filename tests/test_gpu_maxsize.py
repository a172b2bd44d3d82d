"""The largest cube the fused step takes, against the oracle.

Each variable's column of a box holds ktot * jtot * pitch cells (nx + 5 cells per axis with
the ghost and face layers, x rows padded to a multiple of 16 past a 14-cell offset), and the
fused kernels address it with 32-bit byte offsets, so a column must stay under 4 GiB:
798^3 is the largest cube (803 * 803 * 832 cells, 4.29e9 bytes per variable; 180 GB of
device memory with the second state copy and the divergence) and 799^3 is refused.

The problem is the reference KH problem (src_cpp_kokkos_host/main.cpp:188-226; KH/main.cpp
problem generator) with its z-dependent velocity kick switched off (rrv = 0): the state is
then independent of z, every z face flux difference is exactly zero, and every z-plane of
the 798^3 box evolves with the same bits.  So the whole box is checked plane by plane:
every interior plane bitwise equal to the first, and that plane against the oracle run on a
798 x 798 x 4 periodic box of the same cell size (rel L1 <= 1e-10 after 10 steps, dt
sequences equal to 1e-12).  This is the loop body of src_cpp_kokkos_host/main.cpp:229-259 at
the largest size one MI355X runs on the fused path.
"""
import sys

import numpy as np
import pytest

from mhdturbulence_amd import Box, Config
from mhdturbulence_amd.capi import MhdError
from parity import rel_l1

pytestmark = pytest.mark.gpu

N_MAX = 798
STEPS = 10
TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_largest_fused_cube_z_invariant_kh_matches_oracle(oracle_lib):
    n = N_MAX
    dz = 1.0 / n
    o = oracle_lib.Oracle(nx=n, ny=n, nz=4, bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -1.0, -0.5),
                          xmax=(0.5, 1.0, -0.5 + 4 * dz), gam=1.4, cfl=0.05, energy_flux=0, grid_shift=1)
    o.generate_grid()
    o.generate_problem_kh(vy_amp=1.0e-3, rrv=0.0)
    say(f"max size: oracle {n}x{n}x4, {STEPS} steps")
    t_o, log_o = o.run(STEPS)
    ref = o.interior(o.P)  # (nv, 4, n, n)
    for k in range(1, 4):
        assert np.array_equal(ref[:, k], ref[:, 0]), "the oracle's periodic z-invariant box kept its planes equal"

    say(f"max size: creating the {n}^3 box (fused step, FP64)")
    b = Box(Config(nx=n, ny=n, nz=n, bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -1.0, -0.5), xmax=(0.5, 1.0, 0.5),
                   gam=1.4, cfl=0.05, energy_flux=0, grid_shift=1))
    try:
        b.generate_grid()
        b.generate_problem_kh(vy_amp=1.0e-3, rrv=0.0)
        say(f"max size: {n}^3, {STEPS} fused steps")
        t_g, log_g, done = b.run(STEPS, mode=0)
        assert done == STEPS and b.status() == 0
        np.testing.assert_allclose(log_g, log_o, rtol=1e-12)
        assert t_g == pytest.approx(t_o, rel=1e-12)
        say(f"max size: {b.device_bytes() / 2 ** 30:.0f} GiB on the device; downloading P "
            f"({(n + 5) ** 3 * 12 * 8 / 2 ** 30:.0f} GiB on the host)")
        P = b.download_prim()
    finally:
        b.close()
    first = P[2, 2:n + 2, :, 2:n + 2]
    for k in range(3, n + 2):
        if not np.array_equal(P[k, 2:n + 2, :, 2:n + 2], first):
            raise AssertionError(f"plane k={k - 2} differs from plane 0 in a z-invariant state")
    got = np.ascontiguousarray(first.transpose(1, 0, 2))[:, None]  # (nv, 1, n, n)
    del P
    err = rel_l1(got, ref[:, :1])
    say(f"max size: {n}^3 plane vs oracle max relL1 P={err.max():.3e}")
    assert np.all(err <= TOL), err


def test_one_cell_past_the_largest_fused_cube_is_refused():
    n = N_MAX + 1
    say(f"max size: {n}^3 must be refused by the fused step")
    b = Box(Config(nx=n, ny=n, nz=n))
    try:
        b.generate_grid()
        b.generate_problem_kh()
        with pytest.raises(MhdError, match="4 GiB"):
            b.run(1, mode=0)
    finally:
        b.close()
