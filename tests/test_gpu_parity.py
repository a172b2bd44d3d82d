"""Parity of the HIP library (through the C ABI) with the CPU oracle on identical inputs.

Bar: FP64 per-variable relative L1 <= 1e-10 after 100 steps, equal dt sequences, and
each reference step function matching the oracle's to round-off.  Boxes are small
enough for the oracle to finish in seconds.
"""
import sys

import numpy as np
import pytest

from mhdturbulence_amd import Box, Config
import parity

pytestmark = pytest.mark.gpu

KH_BC = (1, 1, 2, 2, 1, 1)
TOL_L1 = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l1(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return parity.rel_l1(a, b, "prim" if a.shape[0] == 12 else "cons")


def pair(orc, n=(32, 32, 32), problem="kh", upload=True, **kw):
    """an oracle box and a GPU box with the same grid and state"""
    o = orc.Oracle(nx=n[0], ny=n[1], nz=n[2], **{k: v for k, v in kw.items() if k != "precision"})
    o.generate_grid()
    if problem == "kh":
        o.generate_problem_kh()
    elif problem == "kh2":
        o.generate_problem_kh(vy_amp=1e-2)
    else:
        o.generate_problem_shear(**problem)
    b = Box(Config(nx=n[0], ny=n[1], nz=n[2], **kw))
    b.set_grid(o.grid(1), o.grid(1, True), o.grid(2), o.grid(2, True), o.grid(3), o.grid(3, True))
    if upload:
        b.upload_prim(o.P)
        b.upload_cons(o.U)
    return o, b


@pytest.fixture(scope="module")
def orc(oracle_lib):
    return oracle_lib


def test_device_grid_is_bitwise_the_oracle_grid(orc):
    for shift in (1, 0):
        say("grid shift", shift)
        o = orc.Oracle(nx=40, ny=24, nz=16, grid_shift=shift)
        o.generate_grid()
        b = Box(Config(nx=40, ny=24, nz=16, grid_shift=shift))
        b.generate_grid()
        got = b.get_grid()
        ref = [o.grid(1), o.grid(1, True), o.grid(2), o.grid(2, True), o.grid(3), o.grid(3, True)]
        for g, r, n in zip(got, ref, (41, 41, 25, 25, 17, 17)):
            np.testing.assert_array_equal(g[:n + 3], r[:n + 3])


def test_upload_download_roundtrip(orc):
    o, b = pair(orc, n=(12, 10, 8))
    np.testing.assert_array_equal(b.download_prim(), o.P)
    np.testing.assert_array_equal(b.download_cons(), o.U)
    # into caller arrays that hold other values: every element is overwritten
    P, U = np.full_like(o.P, np.nan), np.full_like(o.U, -1.0)
    assert b.download_prim(out=P) is P and b.download_cons(out=U) is U
    np.testing.assert_array_equal(P, o.P)
    np.testing.assert_array_equal(U, o.U)
    with pytest.raises(ValueError):
        b.download_prim(out=U)


@pytest.mark.parametrize("problem", ["kh", "shear"])
def test_device_initial_state(orc, problem):
    say("device IC", problem)
    o, b = pair(orc, n=(24, 20, 16), problem="kh" if problem == "kh" else dict(bx0=0.1, a=0.05, seed=7),
                upload=False, gam=5.0 / 3.0 if problem == "shear" else 1.4)
    if problem == "kh":
        b.generate_problem_kh()
    else:
        b.generate_problem_shear(bx0=0.1, a=0.05, seed=7)
    # libm vs device math differ in the last bits of tanh/exp/sin/cos only
    np.testing.assert_allclose(o.interior(b.download_prim()), o.interior(o.P), rtol=0, atol=2e-15)
    np.testing.assert_allclose(o.interior(b.download_cons()), o.interior(o.U), rtol=0, atol=5e-15)


@pytest.mark.parametrize("eflux,problem", [(0, "kh2"), (1, "shear"), (0, "shear")])
def test_reference_functions_one_by_one(orc, eflux, problem):
    say("one-by-one", eflux, problem)
    prob = problem if problem != "shear" else dict(bx0=0.2, a=0.08, amp=0.05)
    o, b = pair(orc, n=(20, 18, 16), problem=prob, energy_flux=eflux)
    dt_o, dt_g = o.control_timestep(), b.control_timestep()
    assert dt_g == pytest.approx(dt_o, rel=1e-15)
    o.set_boundary()
    b.set_boundary_condition()
    Pg = b.download_prim()
    # every ghost the step reads: x ghosts at interior (j,k), y at interior (i,k), z at interior (i,j)
    for sl in (np.s_[2:-3, 2:-3, :, 0:2], np.s_[2:-3, 2:-3, :, -3:-1], np.s_[2:-3, 0:2, :, 2:-3],
               np.s_[2:-3, -3:-1, :, 2:-3], np.s_[0:2, 2:-3, :, 2:-3], np.s_[-3:-1, 2:-3, :, 2:-3]):
        np.testing.assert_array_equal(Pg[sl], o.P[sl])
    ch_o, ch_g = o.evaluate_ch(), b.evaluate_ch()
    assert ch_g == pytest.approx(ch_o, rel=1e-15)
    for d in (1, 2, 3):
        o.numerical_flux(d, ch_o)
        b.get_numerical_flux(d, ch_o)
        Fo, Fg = o.F(d), b.download_flux(d)
        sl = [slice(2, -3)] * 3
        sl[2 - (d - 1)] = slice(2, -2)  # faces run one further along the sweep axis
        s = (sl[0], sl[1], slice(None), sl[2])
        # relative to the largest flux component: a component that is zero up to round-off
        # (the z energy flux of a state with vz = 0 and uniform pressure) has no own scale
        scale = np.abs(Fo[s]).max() + 1e-300
        err = (np.abs(Fg[s] - Fo[s]).max(axis=(0, 1, 3))) / scale
        assert np.all(err < 1e-12), (d, err)
    o.update_conserv(dt_o)
    b.update_conserv_u(dt_o)
    o.damp_psi(dt_o, ch_o)
    b.damp_psi(dt_o, ch_o)
    assert np.all(rel_l1(o.interior(b.download_cons()), o.interior(o.U)) < 1e-14)
    o.update_primitive()
    b.update_primitiv_p()
    assert np.all(rel_l1(o.interior(b.download_prim()), o.interior(o.P)) < 1e-14)


def _wide_state(o, smooth: bool, gam: float):
    """primitive states over many decades, written into the oracle's P (ghosts included).
    smooth=False: each cell drawn independently (density 1e-4..1e3, pressure 1e-3..1e4,
    |v| up to 5, |B| up to 20), for face-by-face flux checks.  smooth=True: density and
    pressure waves of +-3.5 decades across the box with 25 % cell noise, Mach and plasma beta
    of order one, so that a few steps stay physical."""
    rng = np.random.default_rng(20261203)
    P = o.P
    K, J, _, I = P.shape
    k, j, i = np.meshgrid(np.arange(K), np.arange(J), np.arange(I), indexing="ij")
    if smooth:
        rho = 10.0 ** (3.5 * np.sin(2 * np.pi * i / 20) * np.cos(2 * np.pi * k / 16) + 0.1 * rng.uniform(-1, 1, i.shape))
        p = 10.0 ** (3.5 * np.cos(2 * np.pi * j / 18 + 0.3) * np.sin(2 * np.pi * i / 20 + 1.0) +
                     0.1 * rng.uniform(-1, 1, i.shape))
        vs, bs = 0.5 * np.sqrt(gam * p / rho), np.sqrt(p)
    else:
        rho = 10.0 ** rng.uniform(-4, 3, i.shape)
        p = 10.0 ** rng.uniform(-3, 4, i.shape)
        vs, bs = 5.0, 20.0
    P[:, :, 0] = rho
    for c in (1, 2, 3):
        P[:, :, c] = rng.uniform(-1, 1, i.shape) * vs
    P[:, :, 4] = p / ((gam - 1.0) * rho)
    P[:, :, 5] = p
    P[:, :, 6] = np.sqrt(gam * p / rho)
    for c in (7, 8, 9):
        P[:, :, c] = rng.uniform(-1, 1, i.shape) * bs
    P[:, :, 10] = rng.uniform(-1, 1, i.shape)
    P[:, :, 11] = rng.uniform(0, 1, i.shape)
    o.prim_to_cons_all()


def test_wide_range_states(orc):
    """states over many decades: the fluxes of every direction face by face, then three
    fused and three staged steps, against the oracle.  Reaches the HLLD's star-state floors
    and the products its shared reciprocals form across the range the solver meets."""
    say("wide-range states 20x18x16")
    n = (20, 18, 16)
    kw = dict(energy_flux=1, gam=5.0 / 3.0, bc=KH_BC)
    o, b = pair(orc, n=n, problem="kh", upload=False, **kw)
    _wide_state(o, False, kw["gam"])
    b.upload_prim(o.P)
    b.upload_cons(o.U)
    dt_o, dt_g = o.control_timestep(), b.control_timestep()
    assert dt_g == pytest.approx(dt_o, rel=1e-15)
    o.set_boundary()
    b.set_boundary_condition()
    ch_o, ch_g = o.evaluate_ch(), b.evaluate_ch()
    assert ch_g == pytest.approx(ch_o, rel=1e-15)
    for d in (1, 2, 3):
        o.numerical_flux(d, ch_o)
        b.get_numerical_flux(d, ch_o)
        Fo, Fg = o.F(d), b.download_flux(d)
        sl = [slice(2, -3)] * 3
        sl[2 - (d - 1)] = slice(2, -2)
        s = (sl[0], sl[1], slice(None), sl[2])
        # each face against its own largest component: the faces span many decades
        scale = np.abs(Fo[s]).max(axis=2, keepdims=True) + 1e-300
        err = (np.abs(Fg[s] - Fo[s]) / scale).max()
        assert np.isfinite(Fg[s]).all()
        assert err < 1e-11, (d, err)
    b.close()
    for mode in (0, 1):
        o2, b2 = pair(orc, n=n, problem="kh", upload=False, **kw)
        _wide_state(o2, True, kw["gam"])
        b2.upload_prim(o2.P)
        b2.upload_cons(o2.U)
        t_o, log_o = o2.run(3)
        t_g, log_g, done = b2.run(3, mode=mode)
        assert done == 3
        np.testing.assert_allclose(log_g, log_o, rtol=1e-12)
        errP = rel_l1(o2.interior(b2.download_prim()), o2.interior(o2.P))
        errU = rel_l1(o2.interior(b2.download_cons()), o2.interior(o2.U))
        say(f"  mode {mode}: max relL1 P={errP.max():.3e} U={errU.max():.3e}")
        assert np.all(errP < 1e-12) and np.all(errU < 1e-12), (mode, errP, errU)
        b2.close()


def _floor_state(o, gam):
    """states that take the limiter's and the roots' zero cases: half the box uniform
    (every slope pair 0, 0: dp + dm = 0), the other half a checkerboard (every cell a local
    extremum in each direction: dp = -dm, dp + dm = 0 with dp dm < 0), signed zeros in v2
    and psi, and in the uniform half c_s = 1, rho = 1, B = (1, 0, 0), where the fast speed's
    inner root of an x face is sqrt((c_s^2 + B^2/rho)^2 - 4 c_s^2 B_x^2/rho) = sqrt(0)"""
    P = o.P
    K, J, _, I = P.shape
    k, j, i = np.meshgrid(np.arange(K), np.arange(J), np.arange(I), indexing="ij")
    chk = ((i + j + k) % 2).astype(float)
    half = i < I // 2
    rho = np.where(half, 1.0, 1.0 + 0.5 * chk)
    e0 = 1.0 / (gam * (gam - 1.0))
    P[:, :, 0] = rho
    P[:, :, 1] = np.where(half, 0.0, 0.3 * chk - 0.1)
    P[:, :, 2] = np.where(half, -0.0, 0.2 * (1.0 - 2.0 * chk))
    P[:, :, 3] = 0.0
    P[:, :, 4] = e0
    P[:, :, 5] = (gam - 1.0) * rho * e0
    P[:, :, 6] = np.where(half, 1.0, np.sqrt(gam * (gam - 1.0) * e0))
    P[:, :, 7] = np.where(half, 1.0, 0.5 * chk)
    P[:, :, 8] = 0.0
    P[:, :, 9] = 0.0
    P[:, :, 10] = np.where(half, 0.0, -0.0)
    P[:, :, 11] = np.where(half, 0.0, chk)
    o.prim_to_cons_all()


def test_limiter_and_root_zero_cases(orc):
    """the limiter's zero slope sums and opposite-sign slope pairs, signed zeros and a fast
    speed whose inner root is of exactly 0 (the cases the limiter's and the roots' 1e-300
    floors stand in for selects): every face of every direction, then three fused and three
    staged steps, against the oracle"""
    say("limiter and root zero cases 20x18x16")
    n = (20, 18, 16)
    kw = dict(energy_flux=1, gam=5.0 / 3.0, bc=KH_BC)
    o, b = pair(orc, n=n, problem="kh", upload=False, **kw)
    _floor_state(o, kw["gam"])
    b.upload_prim(o.P)
    b.upload_cons(o.U)
    dt_o, dt_g = o.control_timestep(), b.control_timestep()
    assert dt_g == pytest.approx(dt_o, rel=1e-15)
    o.set_boundary()
    b.set_boundary_condition()
    ch_o, ch_g = o.evaluate_ch(), b.evaluate_ch()
    assert ch_g == pytest.approx(ch_o, rel=1e-15)
    for d in (1, 2, 3):
        o.numerical_flux(d, ch_o)
        b.get_numerical_flux(d, ch_o)
        Fo, Fg = o.F(d), b.download_flux(d)
        sl = [slice(2, -3)] * 3
        sl[2 - (d - 1)] = slice(2, -2)
        s = (sl[0], sl[1], slice(None), sl[2])
        assert np.isfinite(Fg[s]).all()
        scale = np.abs(Fo[s]).max(axis=2, keepdims=True) + 1e-300
        err = (np.abs(Fg[s] - Fo[s]) / scale).max()
        assert err < 1e-13, (d, err)
    b.close()
    for mode in (0, 1):
        o2, b2 = pair(orc, n=n, problem="kh", upload=False, **kw)
        _floor_state(o2, kw["gam"])
        b2.upload_prim(o2.P)
        b2.upload_cons(o2.U)
        t_o, log_o = o2.run(3)
        t_g, log_g, done = b2.run(3, mode=mode)
        assert done == 3
        np.testing.assert_allclose(log_g, log_o, rtol=1e-13)
        errP = rel_l1(o2.interior(b2.download_prim()), o2.interior(o2.P))
        errU = rel_l1(o2.interior(b2.download_cons()), o2.interior(o2.U))
        say(f"  mode {mode}: max relL1 P={errP.max():.3e} U={errU.max():.3e}")
        assert np.all(errP < 1e-13) and np.all(errU < 1e-13), (mode, errP, errU)
        b2.close()


def _run_parity(orc, n, steps, mode, problem="kh", **kw):
    o, b = pair(orc, n=n, problem=problem, **kw)
    t_o, log_o = o.run(steps)
    t_g, log_g, done = b.run(steps, mode=mode)
    assert done == steps
    np.testing.assert_allclose(log_g, log_o, rtol=1e-12)
    assert t_g == pytest.approx(t_o, rel=1e-12)
    errP = rel_l1(o.interior(b.download_prim()), o.interior(o.P))
    errU = rel_l1(o.interior(b.download_cons()), o.interior(o.U))
    say(f"n={n} steps={steps} mode={mode} max relL1 P={errP.max():.3e} U={errU.max():.3e}")
    return errP, errU, b, o


@pytest.mark.parametrize("mode", [1, 0])
def test_kh_64cube_100_steps(orc, mode):
    say("KH 64^3 100 steps mode", mode)
    errP, errU, _, _ = _run_parity(orc, (64, 64, 64), 100, mode)
    assert np.all(errP <= TOL_L1), errP
    assert np.all(errU <= TOL_L1), errU


def test_kh_150cube_step_100_checkpoint(orc):
    """the reference KH problem at 150^3 (its README benchmark shape, 32-cell strips with
    ragged edges), fused and staged paths against the oracle at step 100"""
    n = (150, 150, 150)
    o = orc.Oracle(nx=n[0], ny=n[1], nz=n[2])
    o.generate_grid()
    o.generate_problem_kh()
    boxes = []
    for mode in (0, 1):
        b = Box(Config(nx=n[0], ny=n[1], nz=n[2]))
        b.set_grid(o.grid(1), o.grid(1, True), o.grid(2), o.grid(2, True), o.grid(3), o.grid(3, True))
        b.upload_prim(o.P)
        b.upload_cons(o.U)
        boxes.append((mode, b))
    say("KH 150^3: oracle 100 steps")
    t_o, log_o = o.run(100)
    for mode, b in boxes:
        t_g, log_g, done = b.run(100, mode=mode)
        assert done == 100
        np.testing.assert_allclose(log_g, log_o, rtol=1e-12)
        assert t_g == pytest.approx(t_o, rel=1e-12)
        errP = rel_l1(o.interior(b.download_prim()), o.interior(o.P))
        errU = rel_l1(o.interior(b.download_cons()), o.interior(o.U))
        say(f"150^3 mode={mode} max relL1 P={errP.max():.3e} U={errU.max():.3e}")
        assert np.all(errP <= TOL_L1), (mode, errP)
        assert np.all(errU <= TOL_L1), (mode, errU)
        b.close()


@pytest.mark.parametrize("mode", [1, 0])
@pytest.mark.parametrize("eflux", [0, 1])
def test_magnetised_anisotropic(orc, mode, eflux):
    say("magnetised 40x24x16 mode", mode, "eflux", eflux)
    errP, errU, _, _ = _run_parity(orc, (40, 24, 16), 60, mode, problem=dict(bx0=0.15, a=0.05, amp=0.02),
                                   energy_flux=eflux, gam=5.0 / 3.0, xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5),
                                   grid_shift=0)
    assert np.all(errP <= TOL_L1), errP
    assert np.all(errU <= TOL_L1), errU


@pytest.mark.parametrize("bc", [(3, 3, 2, 2, 3, 3), (2, 2, 3, 3, 1, 1)])
def test_other_boundaries(orc, bc):
    say("boundaries", bc)
    for mode in (1, 0):
        errP, errU, _, _ = _run_parity(orc, (24, 24, 20), 40, mode, problem=dict(bx0=0.1, a=0.06, amp=0.05), bc=bc,
                                       energy_flux=1)
        assert np.all(errP <= TOL_L1), (mode, errP)


@pytest.mark.parametrize("shape", [(5, 4, 6), (17, 9, 7), (65, 6, 5), (33, 70, 4)])
def test_small_and_ragged_boxes(orc, shape):
    """one-cell and partial strips, fewer rows / planes than a chunk, several x-strips of
    different widths per wavefront, a y chunk split into several pre-pass batches"""
    say("small/ragged", shape)
    for mode in (0, 1):
        errP, errU, _, _ = _run_parity(orc, shape, 15, mode, problem=dict(bx0=0.12, a=0.08, amp=0.03),
                                       energy_flux=1, gam=5.0 / 3.0, xmin=(-0.5, -0.5, -0.5),
                                       xmax=(0.5, 0.5, 0.5), grid_shift=0)
        assert np.all(errP <= TOL_L1), (mode, errP)
        assert np.all(errU <= TOL_L1), (mode, errU)


@pytest.mark.parametrize("shape,bc", [
    ((256, 10, 8), (1, 1, 2, 2, 1, 1)),   # four 64-cell strips, one block per row, x ends paired (periodic)
    ((256, 10, 8), (3, 3, 2, 2, 1, 1)),   # same blocks, x ends from the pre-pass (outflow)
    ((512, 6, 5), (1, 1, 2, 2, 1, 1)),    # two blocks per row: the edge between them from the pre-pass
    ((128, 8, 10), (1, 1, 2, 2, 1, 1)),   # two strips per row, two rows per block
    ((128, 8, 9), (1, 1, 2, 2, 1, 1)),    # odd row count: rows of a block would straddle y chunks, no sharing
    ((64, 6, 8), (1, 1, 2, 2, 1, 1)),     # one strip per row, four rows per block, periodic ends
    ((64, 6, 8), (3, 3, 2, 2, 1, 1)),     # one strip per row, outflow: nothing to share
])
def test_block_cooperative_strip_edges(orc, shape, bc):
    """x-rows of whole-wave strips whose block neighbours exchange edge cells, interface
    states and face fluxes through LDS (FP32 strips): FP64 against the oracle, and the mixed
    path against the oracle and against the same box run with 32-cell strips (no sharing)"""
    say("block-shared strip edges", shape, bc)
    kw = dict(bc=bc, energy_flux=1, gam=5.0 / 3.0, xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5), grid_shift=0)
    ic = dict(bx0=0.1, a=0.06, amp=0.04)
    errP, errU, _, _ = _run_parity(orc, shape, 12, 0, problem=ic, **kw)
    assert np.all(errP <= TOL_L1), errP
    assert np.all(errU <= TOL_L1), errU
    mixed = []
    for sw in (64, 32):
        o, b = pair(orc, n=shape, problem=ic, precision=1, **kw)
        if sw == 32:
            _, ly, lz = b.get_tiling()
            b.set_tiling(32, min(ly, 32), lz)
        o.run(12)
        b.run(12, mode=0)
        P = b.download_prim()
        err = rel_l1(o.interior(P), o.interior(o.P))
        say("mixed sw", sw, "relL1", err)
        assert err[0] < 1e-6 and err[5] < 1e-6
        mixed.append(o.interior(P))
    # the two runs compute the strip-edge faces in different code (march vs pre-pass), where
    # the compiler may contract FP32 products differently: ulp-level flux differences that
    # stay near 1e-10 after 12 steps; a wrong neighbour cell would show at 1e-6 or above
    assert rel_l1(mixed[0], mixed[1]).max() < 1e-8


@pytest.mark.parametrize("shape,steps,tiling", [
    ((256, 10, 8), 12, None),       # four strips of one x-row per block
    ((128, 8, 10), 12, None),       # two strips per row, two rows per block
    ((64, 6, 8), 12, None),         # one strip per row, four rows per block
    ((64, 40, 24), 40, (64, 16, 24)),   # three y chunks: halo rows between chunks
    ((128, 70, 12), 20, (64, 64, 12)),  # 64-row chunks, the longest, and a 6-row last one
    ((192, 10, 6), 12, None),       # one x-row of three strips per block
    ((512, 12, 4), 12, (64, 5, 4)), # one x-row of eight strips per block, three y chunks
])
def test_fp64_two_waves_per_simd(orc, shape, steps, tiling):
    """The FP64 xy march at two wavefronts per SIMD (strip edges shared in the block, 3-slot window
    ring with row j-1 taken from the carried registers), forced on boxes below the default
    rule's wavefront count, against the oracle; then the one-wavefront kernel on the same box"""
    kw = dict(bc=KH_BC, energy_flux=1, gam=5.0 / 3.0, xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5), grid_shift=0)
    ic = dict(bx0=0.1, a=0.06, amp=0.04)
    for n in (2, 1):
        say("two wavefronts per SIMD" if n == 2 else "one wavefront per SIMD", shape, tiling)
        o, b = pair(orc, n=shape, problem=ic, **kw)
        if tiling and n == 1:
            tiling = (tiling[0], min(tiling[1], 46), tiling[2])  # the one-wavefront kernel's LDS limit
        b.set_xy_waves_per_simd(n)
        if tiling:
            b.set_tiling(*tiling)
        assert b.get_xy_waves_per_simd() == n
        assert b.get_tiling()[0] == 64
        t_o, log_o = o.run(steps)
        t_g, log_g, done = b.run(steps, mode=0)
        assert done == steps
        np.testing.assert_allclose(log_g, log_o, rtol=1e-12)
        errP = rel_l1(o.interior(b.download_prim()), o.interior(o.P))
        errU = rel_l1(o.interior(b.download_cons()), o.interior(o.U))
        say(f"  max relL1 P={errP.max():.3e} U={errU.max():.3e}")
        assert np.all(errP <= TOL_L1), errP
        assert np.all(errU <= TOL_L1), errU
        b.close()


def test_two_waves_per_simd_refused_where_the_box_does_not_allow_it(orc):
    """the switch refuses boxes whose strip edges are not all inside a block, and a
    one-wavefront kernel whose y chunk would exceed its LDS; mixed precision reports one"""
    from mhdturbulence_amd.capi import MhdError
    base = dict(energy_flux=1, gam=5.0 / 3.0)
    for shape, bc in (((80, 8, 8), KH_BC), ((256, 8, 8), (3, 3, 2, 2, 1, 1)), ((128, 8, 9), KH_BC),
                      ((320, 8, 8), KH_BC), ((384, 8, 4), KH_BC)):
        b = Box(Config(nx=shape[0], ny=shape[1], nz=shape[2], bc=bc, **base))
        with pytest.raises(MhdError):
            b.set_xy_waves_per_simd(2)
        assert b.get_xy_waves_per_simd() == 1
        b.close()
    b = Box(Config(nx=256, ny=64, nz=8, bc=KH_BC, **base))
    b.set_xy_waves_per_simd(2)
    b.set_tiling(64, 64, 8)
    with pytest.raises(MhdError):
        b.set_xy_waves_per_simd(1)
    assert b.get_xy_waves_per_simd() == 2
    b.set_tiling(32, 16, 8)
    assert b.get_xy_waves_per_simd() == 1   # 32-cell strips: the one-wavefront kernel
    b.close()
    b = Box(Config(nx=256, ny=16, nz=8, bc=KH_BC, precision=2, **base))
    b.set_xy_waves_per_simd(2)
    assert b.get_xy_waves_per_simd() == 1
    b.close()


@pytest.mark.parametrize("variant", ["two_waves", "one_wave", "staged"])
def test_step_on_nonuniform_grid(orc, variant):
    """The whole step on a stretched grid (random cell widths, a different draw per axis):
    every kernel takes each cell's own 1/dx, 1/dy, 1/dz and CFL widths.  The two-wavefront
    march runs 64-row y chunks (64, 64 and 12 rows), whose rows take their 1/dy from one lane
    each; on a uniform grid a wrong row or lane there would not show."""
    import divb_fields as dvf
    n = (256, 140, 6)
    kw = dict(bc=KH_BC, energy_flux=1, gam=5.0 / 3.0, xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5), grid_shift=0)
    say("stretched grid", n, variant)
    o = orc.Oracle(nx=n[0], ny=n[1], nz=n[2], **kw)
    o.generate_grid()
    for ax, seed in enumerate((11, 12, 13)):
        f = dvf.nonuniform_faces(n[ax], -0.5, 1.0, seed)
        o.grid(ax + 1)[:] = f
        o.grid(ax + 1, True)[:] = dvf.centres(f)
    o.generate_problem_shear(bx0=0.1, a=0.06, amp=0.04)
    b = Box(Config(nx=n[0], ny=n[1], nz=n[2], **kw))
    b.set_grid(o.grid(1), o.grid(1, True), o.grid(2), o.grid(2, True), o.grid(3), o.grid(3, True))
    b.upload_prim(o.P)
    b.upload_cons(o.U)
    if variant == "two_waves":
        b.set_xy_waves_per_simd(2)
        b.set_tiling(64, 64, 6)
        assert b.get_tiling() == (64, 64, 6) and b.get_xy_waves_per_simd() == 2
    elif variant == "one_wave":
        b.set_xy_waves_per_simd(1)
    t_o, log_o = o.run(15)
    t_g, log_g, done = b.run(15, mode=1 if variant == "staged" else 0)
    assert done == 15
    np.testing.assert_allclose(log_g, log_o, rtol=1e-12)
    errP = rel_l1(o.interior(b.download_prim()), o.interior(o.P))
    errU = rel_l1(o.interior(b.download_cons()), o.interior(o.U))
    say(f"  max relL1 P={errP.max():.3e} U={errU.max():.3e}")
    assert np.all(errP <= TOL_L1), errP
    assert np.all(errU <= TOL_L1), errU


def test_tmax_stop_matches_reference_loop(orc):
    o, b = pair(orc, n=(16, 16, 16))
    dt0 = o.control_timestep()
    tmax = 5.5 * dt0
    t, log, done = b.run(50, tmax=tmax, mode=0)
    # the reference loop breaks after the step that takes time past time_max
    assert done == 6 and t > tmax and t - log[-1] <= tmax
    t_o, log_o = o.run(done)
    np.testing.assert_allclose(log, log_o, rtol=1e-12)


def test_run_continues_across_calls(orc):
    o, b = pair(orc, n=(24, 24, 24))
    t1, l1, _ = b.run(7, mode=0)
    t2, l2, _ = b.run(8, time=t1, mode=0)
    t_o, log_o = o.run(15)
    np.testing.assert_allclose(np.concatenate([l1, l2]), log_o, rtol=1e-12)
    assert np.all(rel_l1(o.interior(b.download_cons()), o.interior(o.U)) <= TOL_L1)


@pytest.mark.parametrize("between", ["nothing", "download", "status", "tiling", "staged"])
def test_fused_run_continues_from_live_state(orc, between):
    """a fused run leaves the state live in its buffers (after 3 steps: in the second one);
    the next fused run continues from there, or, after a call that settles the state, starts
    again from P and U; either way 3 + 4 steps equal the oracle's 7"""
    o, b = pair(orc, n=(24, 20, 16), problem=dict(bx0=0.1), gam=5.0 / 3.0, energy_flux=1)
    t1, l1, _ = b.run(3, mode=0)
    if between == "download":
        b.download_prim()
    elif between == "status":
        assert b.status() == 0
    elif between == "tiling":
        b.set_tiling(*b.get_tiling())
    t2, l2, _ = b.run(4, time=t1, mode=1 if between == "staged" else 0)
    t_o, log_o = o.run(7)
    np.testing.assert_allclose(np.concatenate([l1, l2]), log_o, rtol=1e-12)
    assert t2 == pytest.approx(t_o, rel=1e-12)
    errP = rel_l1(o.interior(b.download_prim()), o.interior(o.P))
    errU = rel_l1(o.interior(b.download_cons()), o.interior(o.U))
    say(f"between {between}: max relL1 P={errP.max():.3e} U={errU.max():.3e}")
    assert np.all(errP <= TOL_L1) and np.all(errU <= TOL_L1)


@pytest.mark.parametrize("stop", [False, True])
def test_fused_step_pair_after_run(orc, stop):
    """the one-step fused pair (the C++ drop-in's recorded step) after a fused run: one that
    ran all its steps (its state left live, 3 steps: in the second buffer) or one that
    stopped past tmax; two more steps through the pair equal the oracle's continuation"""
    import ctypes as C
    o, b = pair(orc, n=(24, 20, 16), problem=dict(bx0=0.1), gam=5.0 / 3.0, energy_flux=1)
    tmax = 0.0
    if stop:
        tmax = 2.5 * o.control_timestep()
    t, log, done = b.run(50 if stop else 3, tmax=tmax, mode=0)
    assert done == 3
    dts = list(log)
    for _ in range(2):
        dt, ch, st = C.c_double(), C.c_double(), C.c_int()
        assert b.L.mhd_fused_begin_step(b.h, C.byref(dt), C.byref(ch), C.byref(st)) == 0 and st.value == 0
        assert b.L.mhd_fused_step(b.h) == 0
        dts.append(dt.value)
    _, log_o = o.run(5)
    np.testing.assert_allclose(dts, log_o, rtol=1e-12)
    errU = rel_l1(o.interior(b.download_cons()), o.interior(o.U))
    say(f"stop {stop}: max relL1 U={errU.max():.3e}")
    assert np.all(errU <= TOL_L1)


@pytest.mark.parametrize("mode,steps", [(0, 3), (0, 4), (1, 3)])
def test_loop_leaves_ghost_cells_of_u(orc, mode, steps):
    """The reference's loop writes only interior cells of U (src_cpp_kokkos_host/mhd.cpp:1431,
    1625; SetBoundaryCondition fills P): after a run, U's ghost cells hold what was uploaded,
    bitwise, as the oracle's do, after an odd number of fused steps too (the state then
    comes back from the second buffer, interior only)"""
    o, b = pair(orc, n=(24, 16, 12), problem=dict(bx0=0.1), gam=5.0 / 3.0, energy_flux=1)
    U0 = o.U.copy()
    idx = np.arange(U0.size).reshape(U0.shape)
    ghost = np.ones(U0.size, bool)
    ghost[o.interior(idx).ravel()] = False
    ghost = ghost.reshape(U0.shape)
    # ghost values no step computes, so a buffer's stale or zeroed ghosts cannot pass
    U_up = U0.copy()
    U_up[ghost] = 7.0 + 1e-3 * np.arange(ghost.sum())
    b.upload_cons(U_up)
    b.run(steps, mode=mode)
    o.run(steps)
    U = b.download_cons()
    say(f"mode {mode} steps {steps}: {ghost.sum()} ghost values")
    assert np.array_equal(o.U[ghost], U0[ghost])
    assert np.array_equal(U[ghost], U_up[ghost])
    assert np.all(rel_l1(o.interior(U), o.interior(o.U)) <= TOL_L1)


def test_diagnostics_match(orc):
    o, b = pair(orc, n=(32, 32, 16), problem=dict(bx0=0.1, a=0.05, amp=0.03), energy_flux=1)
    o.run(20)
    b.run(20, mode=0)
    do, dg = o.diagnostics(), b.diagnostics()
    np.testing.assert_allclose(dg[:3], do[:3], rtol=1e-12)
    assert dg[3] == pytest.approx(do[3], rel=1e-12)


@pytest.mark.parametrize("kind", ["linear", "quadratic", "quadratic_low", "sine", "solenoidal"])
def test_divb_closed_form_on_device(orc, kind):
    """mhd_diagnostics of fields whose discrete divergence is known in closed form
    (tests/divb_fields.py), on non-uniform anisotropic grids uploaded with mhd_set_grid"""
    import divb_fields as dvf
    from test_oracle import _divb_box
    n = (14, 11, 9)
    o = _divb_box(orc, kind, n)
    b = Box(Config(nx=n[0], ny=n[1], nz=n[2], bc=(1, 1, 1, 1, 1, 1), xmin=(0.3, 0.2, 0.1), xmax=(1.3, 1.0, 0.7),
                   grid_shift=0))
    b.set_grid(o.grid(1), o.grid(1, True), o.grid(2), o.grid(2, True), o.grid(3), o.grid(3, True))
    b.upload_prim(o.P)
    dg, do = b.diagnostics(), o.diagnostics()
    want = dvf.expected_max_divb(kind, o.grid(1, True), o.grid(2, True), o.grid(3, True), n)
    if kind == "solenoidal":
        assert dg[3] == 0.0
    else:
        assert dg[3] == pytest.approx(want, rel=1e-13), (kind, dg[3], want)
    np.testing.assert_allclose(dg[:3], do[:3], rtol=1e-13)


@pytest.mark.parametrize("precision", [1, 2])
def test_mixed_precision_close_to_fp64(orc, precision):
    """precision 1: FP32 reconstruction + HLLD; 2: also the x-y divergence stored in FP32
    between the fused kernels (the update itself in FP64)"""
    say("mixed precision", precision)
    o, b = pair(orc, n=(32, 32, 32), problem=dict(bx0=0.1, a=0.05, amp=0.02), energy_flux=1, gam=5.0 / 3.0,
                precision=precision)
    o.run(50)
    b.run(50, mode=0)
    err = rel_l1(o.interior(b.download_prim()), o.interior(o.P))
    say("mixed relL1", err)
    assert err[0] < 1e-6 and err[5] < 1e-6
    assert np.all(err[[0, 1, 5, 7, 11]] < 1e-4)


def test_kernel_profile_runs(orc):
    o, b = pair(orc, n=(32, 32, 32))
    for mode in (1, 0):
        prof = b.profile_kernels(3, mode=mode)
        assert len(prof) >= 2 and all(v >= 0 for v in prof.values())


def test_kh64_tprof_long_run_matches_oracle_fixture(tmp_path):
    """12,700 steps of the reference KH problem at 64^3 (t = 0 to 4.5): every t-prof row of
    the fused FP64 path equals the oracle's (tests/golden/kh64_tprof.csv, written by
    tests/golden/make_kh64_tprof.py with the same sampling) to round-off."""
    import os
    from mhdturbulence_amd import tprof
    gold = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "kh64_tprof.csv"), comments="#")
    r = tprof.run(64, 4.5, 0.05, str(tmp_path))
    got = np.loadtxt(tmp_path / "t-prof.csv", comments="#")
    assert r["steps"] == int(gold[-1, 0]) and got.shape[0] == gold.shape[0]
    np.testing.assert_allclose(got[:, 0], gold[:, 1], rtol=1e-12, atol=1e-15)  # sample times
    np.testing.assert_allclose(got[:, 1], gold[:, 2], rtol=1e-10)  # mix
    np.testing.assert_allclose(got[:, 2], gold[:, 3], rtol=1e-10)  # sqrt(<vy^2>)
