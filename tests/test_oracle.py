"""Pin the CPU oracle (oracle/mhd_oracle.cpp) against what fixes it independently.

The reference ships no tests and cannot be compiled here, so each function of the
restatement is checked against: golden values the reference produced (the t-prof
first row recorded from a src_cpp_omp_device run, the analysis/vissimple t=0
snapshot), closed forms, textbook Riemann solvers written from the papers, exact
invariants (conservation, mirror symmetry, consistency), brute-force index checks
and the convergence order of the whole scheme.
"""
import os

import numpy as np
import pytest

import mhd_textbook as tb

pytestmark = []
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def orc(oracle_lib):
    return oracle_lib


# --------------------------------------------------------------------------- van Leer
def test_vanleer_closed_form(orc):
    rng = np.random.default_rng(1)
    for _ in range(2000):
        a, b = rng.normal(size=2) * 10.0 ** rng.integers(-8, 3, size=2)
        got = orc.vanleer(a, b)
        if a * b > 0:
            assert got == pytest.approx(2 * a * b / (a + b), rel=1e-15)
            assert min(abs(a), abs(b)) <= abs(got) <= 2 * min(abs(a), abs(b)) * (1 + 1e-15)  # TVD bound
            assert np.sign(got) == np.sign(a)
        else:
            assert got == 0.0
    assert orc.vanleer(0.3, 0.3) == 0.3
    assert orc.vanleer(0.0, 1.0) == 0.0
    assert orc.vanleer(-1.0, 2.0) == 0.0


# --------------------------------------------------------------------------- Prim -> Riemann vector
@pytest.mark.parametrize("d", [1, 2, 3])
def test_prim2riemann_matches_textbook(orc, d):
    rng = np.random.default_rng(10 + d)
    for _ in range(200):
        q = tb.random_prim(rng)
        c = orc.prim2riemann(q, d, energy_flux=1)
        U = tb.conserved(q)
        F = tb.physical_flux(q, d - 1)
        dn = d - 1
        t1, t2 = (dn + 1) % 3, (dn + 2) % 3
        rot = [0, 1 + dn, 1 + t1, 1 + t2, 4, 5 + dn, 5 + t1, 5 + t2, 8, 9]
        np.testing.assert_allclose(c[0:10], U[rot], rtol=1e-14, atol=1e-15)
        Fr = F[rot]
        Fr[5] = 0.0
        Fr[8] = 0.0
        np.testing.assert_allclose(c[10:20], Fr, rtol=1e-13, atol=1e-14)
        assert c[20] == pytest.approx(tb.fast_speed(q, dn), rel=1e-14)
        assert c[21] == q[1 + dn]
        assert c[22] == pytest.approx(q[5] + 0.5 * q[7:10] @ q[7:10], rel=1e-15)
        # energy_flux=0 differs from the physical flux by exactly (|B|^2 - v.B) B_n in the energy entry
        c0 = orc.prim2riemann(q, d, energy_flux=0)
        np.testing.assert_array_equal(np.delete(c0, 14), np.delete(c, 14))
        b = q[7:10]
        assert c0[14] - c[14] == pytest.approx(-(b @ b - q[1:4] @ b) * b[dn], rel=1e-12, abs=1e-14)


# --------------------------------------------------------------------------- HLLD
def _hlld_from_prims(orc, qL, qR, d, eflux=1):
    """HLLD output (rotated frame, KH/mhd.cpp:393-415 slots) turned back into Cartesian order."""
    nf = orc.hlld(orc.prim2riemann(qL, d, eflux), orc.prim2riemann(qR, d, eflux))
    dn = d - 1
    out = nf.copy()
    for c in range(3):
        s = (c - dn) % 3
        out[1 + c] = nf[1 + s]
        out[5 + c] = nf[5 + s]
    return out


@pytest.mark.parametrize("d", [1, 2, 3])
def test_hlld_consistency(orc, d):
    """HLLD(U, U) = F(U) for any single state."""
    rng = np.random.default_rng(20 + d)
    for _ in range(300):
        q = tb.random_prim(rng, vscale=3.0)
        nf = _hlld_from_prims(orc, q, q, d)
        F = tb.physical_flux(q, d - 1)
        mask = np.ones(10, bool)
        mask[[4 + d, 8]] = False  # B_n and psi entries come from GLM
        np.testing.assert_allclose(nf[mask], F[mask], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_hlld_matches_textbook_all_regions(orc, d):
    """Branch-free HLLD of the restatement == Miyoshi-Kusano HLLD written with explicit branches."""
    rng = np.random.default_rng(30 + d)
    seen = set()
    for n in range(3000):
        vs = 3.0 if n % 3 == 0 else 0.7
        qL = tb.random_prim(rng, vscale=vs)
        qR = tb.random_prim(rng, vscale=vs)
        qR[6 + d] = qL[6 + d]  # continuous normal field
        ref, region = tb.hlld(qL, qR, d - 1)
        seen.add(region)
        got = _hlld_from_prims(orc, qL, qR, d)
        mask = np.ones(10, bool)
        mask[[4 + d, 8]] = False
        scale = np.max(np.abs(ref)) + 1.0
        np.testing.assert_allclose(got[mask], ref[mask], rtol=0, atol=2e-12 * scale, err_msg=region)
    assert seen == {"L", "L*", "L**", "R**", "R*", "R"}


@pytest.mark.parametrize("d", [1, 2, 3])
def test_hlld_without_field_is_hllc(orc, d):
    rng = np.random.default_rng(40 + d)
    for _ in range(1000):
        qL = tb.random_prim(rng, bscale=0.0, vscale=2.0)
        qR = tb.random_prim(rng, bscale=0.0, vscale=2.0)
        got = _hlld_from_prims(orc, qL, qR, d)
        ref = tb.hllc(qL, qR, d - 1)
        mask = np.ones(10, bool)
        mask[[4 + d, 8]] = False
        np.testing.assert_allclose(got[mask], ref[mask], rtol=0, atol=1e-12 * (np.max(np.abs(ref)) + 1.0))


def test_hlld_upwinding_exact(orc):
    """supersonic to the right: the flux is the left physical flux; to the left: the right one."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        qL = tb.random_prim(rng, bscale=0.2, vscale=0.1)
        qR = tb.random_prim(rng, bscale=0.2, vscale=0.1)
        qR[7] = qL[7]
        for sgn in (1, -1):
            a, b = qL.copy(), qR.copy()
            a[1] += sgn * 20.0
            b[1] += sgn * 20.0
            got = _hlld_from_prims(orc, a, b, 1)
            ref = tb.physical_flux(a if sgn > 0 else b, 0)
            np.testing.assert_allclose(np.delete(got, [5, 8]), np.delete(ref, [5, 8]), rtol=1e-13, atol=1e-12)


# --------------------------------------------------------------------------- face flux (reconstruction + GLM)
def test_face_flux_uniform_state_and_glm(orc):
    rng = np.random.default_rng(6)
    for d in (1, 2, 3):
        q = tb.random_prim(rng)
        chg = 2.5
        f = orc.face_flux(q, q, q, q, d, chg, energy_flux=1)
        F = tb.physical_flux(q, d - 1)
        mask = np.ones(10, bool)
        mask[[4 + d, 8]] = False
        np.testing.assert_allclose(f[mask], F[mask], rtol=1e-12, atol=1e-13)
        # GLM: F(B_n) = psi, F(psi) = c_h^2 B_n for a uniform state
        assert f[4 + d] == pytest.approx(q[10], rel=1e-15)
        assert f[8] == pytest.approx(chg * chg * q[6 + d], rel=1e-14)


def test_face_flux_glm_jump_closed_form(orc):
    """GLM flux of a (B_n, psi) jump with everything else uniform."""
    base = tb.random_prim(np.random.default_rng(7), bscale=0.0)
    qm, qp = base.copy(), base.copy()
    qm[7], qm[10] = 0.3, 0.05
    qp[7], qp[10] = 0.5, -0.02
    chg = 1.7
    # cells f-2, f-1 = qm; f, f+1 = qp -> limited slopes vanish, face states are qm | qp
    f = orc.face_flux(qm, qm, qp, qp, 1, chg, energy_flux=1)
    assert f[5] == pytest.approx(0.5 * (qm[10] + qp[10]) - 0.5 * chg * (qp[7] - qm[7]), rel=1e-14)
    assert f[8] == pytest.approx((0.5 * (qm[7] + qp[7]) - 0.5 / chg * (qp[10] - qm[10])) * chg * chg, rel=1e-14)


def test_face_flux_reconstruction_is_muscl_vanleer(orc):
    """Left/right states are cell value +- half the van Leer slope (checked through a B=0 HLLC face)."""
    rng = np.random.default_rng(8)
    for _ in range(200):
        cells = [tb.random_prim(rng, bscale=0.0, vscale=0.5) for _ in range(4)]
        for c in cells:  # keep e, cs consistent after reconstruction: vary rho, v, X only
            c[4], c[5], c[6] = 2.5, 1.0, np.sqrt(1.4)
            c[0] = 1.0
        got = orc.face_flux(*cells, 1, 1.0, energy_flux=1)

        def vl(a, b):
            return 2 * a * b / (a + b) if a * b > 0 else 0.0
        ql = np.array([cells[1][n] + 0.5 * vl(cells[2][n] - cells[1][n], cells[1][n] - cells[0][n]) for n in range(12)])
        qr = np.array([cells[2][n] - 0.5 * vl(cells[3][n] - cells[2][n], cells[2][n] - cells[1][n]) for n in range(12)])
        ref = tb.hllc(ql, qr, 0)
        mask = np.ones(10, bool)
        mask[[5, 8]] = False
        np.testing.assert_allclose(got[mask], ref[mask], rtol=0, atol=1e-12)


# --------------------------------------------------------------------------- box-level functions
def _box(orc, n=(12, 10, 8), **kw):
    o = orc.Oracle(nx=n[0], ny=n[1], nz=n[2], **kw)
    o.generate_grid()
    return o


def test_grid_formula_and_shift(orc):
    o = _box(orc, n=(64, 64, 64))
    x1a, x1b, x2b = o.grid(1), o.grid(1, True), o.grid(2, True)
    dx = 1.0 / 64
    assert x1a[2] == pytest.approx(-0.5 - dx)       # first interior face sits at xmin - dx
    assert x1b[2] == pytest.approx(-0.5078125)      # = -0.5 - dx/2 (64^3 probe of the reference)
    assert x2b[2] == pytest.approx(-1.015625)
    assert x2b[o.jtot - 4] == pytest.approx(0.953125)
    o0 = _box(orc, n=(64, 64, 64), grid_shift=0)
    assert o0.grid(1)[2] == -0.5


def test_tprof_first_row_matches_reference_run(orc):
    """src_cpp_omp_device at 300x300x10 wrote t-prof row 0: mix = 2.4999999946851207E-02 and
    sqrt(<vy^2>) = 3.5402237636257437E-03 (its initial state: vy amplitude 1e-2)."""
    o = orc.Oracle(nx=300, ny=300, nz=10)
    o.generate_grid()
    o.generate_problem_kh(vy_amp=1e-2)
    d = o.diagnostics()
    assert d[0] == pytest.approx(2.0, rel=1e-11)  # 9e5 cell volumes summed
    assert d[1] == pytest.approx(2.4999999946851207e-02, rel=1e-11)
    assert d[2] == pytest.approx(3.5402237636257437e-03, rel=1e-11)


def test_initial_state_matches_vissimple_snapshot(orc):
    """analysis/vissimple t=0 slice (150^3, unshifted Fortran grid, vy amplitude 1e-3), 3 digits."""
    rows = np.loadtxt(os.path.join(HERE, "golden", "vissimple_t0.csv"))
    o = orc.Oracle(nx=150, ny=150, nz=4, grid_shift=0)
    o.generate_grid()
    o.generate_problem_kh(vy_amp=1e-3, rrv=0.0)
    x1b, x2b = o.grid(1, True), o.grid(2, True)
    P = o.P
    for x, y, d, vy, p, X in rows:
        i = int(np.argmin(np.abs(x1b - x)))
        j = int(np.argmin(np.abs(x2b - y)))
        assert abs(x1b[i] - x) < 5e-3 and abs(x2b[j] - y) < 5e-3
        q = P[2, j, :, i]
        for got, ref in ((q[0], d), (q[2], vy), (q[5], p), (q[11], X)):
            assert got == pytest.approx(ref, rel=5.1e-3, abs=1e-300), (x, y, got, ref)


def _divb_box(orc, kind, n=(14, 11, 9)):
    """an oracle box holding divb_fields' field `kind`: non-uniform grids (a different one per
    axis) for the polynomial and solenoidal fields, the generated uniform grid for the sine"""
    import divb_fields as dvf
    o = orc.Oracle(nx=n[0], ny=n[1], nz=n[2], bc=(1, 1, 1, 1, 1, 1), xmin=(0.3, 0.2, 0.1), xmax=(1.3, 1.0, 0.7),
                   grid_shift=0)
    o.generate_grid()
    if kind != "sine":
        for ax, (lo, ln, seed) in enumerate(((0.3, 1.0, 1), (0.2, 0.8, 2), (0.1, 0.6, 3))):
            f = dvf.nonuniform_faces(n[ax], lo, ln, seed)
            o.grid(ax + 1)[:] = f
            o.grid(ax + 1, True)[:] = dvf.centres(f)
    dvf.fill(o.P, o.grid(1, True), o.grid(2, True), o.grid(3, True), kind)
    return o


@pytest.mark.parametrize("kind", ["linear", "quadratic", "quadratic_low", "sine", "solenoidal"])
def test_divb_closed_form(orc, kind):
    """max|divB| of the oracle's diagnostics against the closed-form central difference of
    fields whose discrete divergence is known (tests/divb_fields.py), on non-uniform and
    anisotropic grids; the volume is the domain's and the sums are those of the cells"""
    import divb_fields as dvf
    n = (14, 11, 9)
    o = _divb_box(orc, kind, n)
    d = o.diagnostics()
    want = dvf.expected_max_divb(kind, o.grid(1, True), o.grid(2, True), o.grid(3, True), n)
    if kind == "solenoidal":
        assert d[3] == 0.0
    else:
        assert d[3] == pytest.approx(want, rel=1e-13), (kind, d[3], want)
    assert d[0] == pytest.approx(1.0 * 0.8 * 0.6, rel=1e-13)
    # the sums are over interior cells, weighted by cell volume
    xa, ya, za = o.grid(1), o.grid(2), o.grid(3)
    dv = (np.diff(za)[2:2 + n[2], None, None] * np.diff(ya)[None, 2:2 + n[1], None] *
          np.diff(xa)[None, None, 2:2 + n[0]])
    Pi = o.interior(o.P)
    X, vy = Pi[11], Pi[2]
    assert d[1] == pytest.approx((X * (1 - X) * dv).sum() / dv.sum(), rel=1e-13)
    assert d[2] == pytest.approx(np.sqrt((vy * vy * dv).sum() / dv.sum()), rel=1e-13)


def test_dt_and_ch_closed_form_uniform(orc):
    o = _box(orc, n=(8, 8, 8), xmin=(0, 0, 0), xmax=(1, 2, 4))
    P = o.P
    rho, v, p, b, gam = 1.3, np.array([0.2, -0.7, 0.1]), 0.8, np.array([0.3, -0.4, 0.5]), 1.4
    P[:, :, 0] = rho
    P[:, :, 1:4] = v[None, None, :, None]
    P[:, :, 5] = p
    P[:, :, 4] = p / ((gam - 1) * rho)
    P[:, :, 6] = np.sqrt(gam * p / rho)
    P[:, :, 7:10] = b[None, None, :, None]
    cs2 = gam * p / rho
    ctot = np.sqrt(cs2 + b @ b / rho)
    dxs = np.array([1 / 8, 2 / 8, 4 / 8])
    assert o.control_timestep() == pytest.approx(0.05 * np.min(dxs / (np.abs(v) + ctot)), rel=1e-14)
    cts = cs2 + b @ b / rho
    cf = np.sqrt(0.5 * (cts + np.sqrt(cts ** 2 - 4 * cs2 * b ** 2 / rho)))
    assert o.evaluate_ch() == pytest.approx(np.max(np.abs(v) + cf), rel=1e-14)


@pytest.mark.parametrize("bc", [(1, 1, 2, 2, 1, 1), (2, 3, 1, 1, 3, 2), (3, 2, 2, 3, 1, 1)])
def test_boundary_brute_force(orc, bc):
    o = _box(orc, n=(6, 5, 4), bc=bc)
    rng = np.random.default_rng(11)
    P = o.P
    P[...] = rng.normal(size=P.shape)
    before = P.copy()
    o.set_boundary()
    nx, ny, nz = 6, 5, 4
    lo = 2
    his = (lo + nx - 1, lo + ny - 1, lo + nz - 1)

    def src(axis, side, g):
        b = bc[2 * axis + side]
        hi = his[axis]
        if b == 1:
            return (lo + g, False) if side else (hi - 1 + g, False)
        if b == 2:
            return (hi - g, True) if side else (lo + 1 - g, True)
        return (hi, False) if side else (lo, False)

    for axis in range(3):
        for side in (0, 1):
            for g in (0, 1):
                s, neg = src(axis, side, g)
                dst = (his[axis] + 1 + g) if side else g
                for a in range(lo, lo + (ny if axis == 0 else nx)):
                    for b_ in range(lo, lo + (nz if axis != 2 else ny)):
                        if axis == 0:
                            got, exp = P[b_, a, :, dst], before[b_, a, :, s].copy()
                        elif axis == 1:
                            got, exp = P[b_, dst, :, a], before[b_, s, :, a].copy()
                        else:
                            got, exp = P[dst, b_, :, a], before[s, b_, :, a].copy()
                        if neg:
                            exp[1 + axis] = -exp[1 + axis]
                        np.testing.assert_array_equal(got, exp)
    # interior untouched
    np.testing.assert_array_equal(P[2:2 + nz, 2:2 + ny, :, 2:2 + nx], before[2:2 + nz, 2:2 + ny, :, 2:2 + nx])


def _total(o, arr):
    x1a, x2a, x3a = o.grid(1), o.grid(2), o.grid(3)
    dv = np.einsum("k,j,i->kji", np.diff(x3a)[2:o.ktot - 3], np.diff(x2a)[2:o.jtot - 3], np.diff(x1a)[2:o.itot - 3])
    return np.einsum("kjni,kji->n", arr[2:o.ktot - 3, 2:o.jtot - 3, :, 2:o.itot - 3], dv)


@pytest.mark.parametrize("eflux", [0, 1])
def test_conservation_periodic(orc, eflux):
    """All-periodic box with B != 0: mass, momentum, energy, B and rhoX totals are conserved."""
    o = _box(orc, n=(16, 12, 10), bc=(1, 1, 1, 1, 1, 1), xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5),
             energy_flux=eflux)
    o.generate_problem_shear(bx0=0.3, amp=0.05, a=0.1)
    o.U[...]  # noqa
    tot0 = _total(o, o.U)
    o.run(20)
    tot1 = _total(o, o.U)
    for n in (0, 1, 2, 3, 4, 5, 6, 7, 9):
        assert abs(tot1[n] - tot0[n]) <= 1e-13 * (abs(tot0[n]) + 1.0), (n, tot0[n], tot1[n])


def test_mirror_symmetry_y(orc):
    """A state symmetric under the mirror y -> -y stays symmetric with reflecting y walls.

    v is a polar vector (vy odd) and B a pseudovector (Bx, Bz odd, By even), so the normal
    field is continuous across the mirror plane; psi is then odd.  The (v.B) B_n energy
    flux respects this symmetry (the |B|^2 B_n form does not)."""
    o = _box(orc, n=(8, 16, 6), bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -1.0, -0.5), xmax=(0.5, 1.0, 0.5), grid_shift=0,
             energy_flux=1)
    rng = np.random.default_rng(12)
    P = o.P
    ny = 16
    half = rng.uniform(0.5, 1.5, size=(o.ktot, ny // 2, o.itot))
    P[:, :, :, :] = 0.0
    for j in range(ny // 2):
        jm, jp = 2 + j, 2 + ny - 1 - j
        for n, val in ((0, half[:, j]), (5, 1.0 + 0.1 * half[:, j]), (1, 0.3 * half[:, j]), (2, 0.1 * half[:, j]),
                       (7, 0.2 * half[:, j]), (8, 0.05 * half[:, j]), (9, 0.1 * half[:, j]),
                       (11, 0.5 * half[:, j])):
            odd = n in (2, 7, 9, 10)
            P[:, jm, n, :] = val
            P[:, jp, n, :] = -val if odd else val
    P[:, :, 4] = P[:, :, 5] / (0.4 * np.where(P[:, :, 0] > 0, P[:, :, 0], 1.0))
    P[:, :, 6] = np.sqrt(1.4 * P[:, :, 5] / np.where(P[:, :, 0] > 0, P[:, :, 0], 1.0))
    o.prim_to_cons_all()
    o.run(5)
    Pi = o.interior(o.P)
    flip = Pi[:, :, ::-1, :].copy()
    flip[[2, 7, 9, 10]] *= -1
    np.testing.assert_allclose(Pi, flip, rtol=0, atol=1e-13)


def _rk2_step(o, dt, ch):
    """one two-stage SSP Runge-Kutta step from the oracle's reference functions:
    U* = E(U^n), U^{n+1} = (U^n + E(U*)) / 2, E = the reference's forward-Euler stage"""
    o.save_cons()
    for stage in (0, 1):
        if stage:
            o.set_boundary()
        for d in (1, 2, 3):
            o.numerical_flux(d, ch)
        o.update_conserv(dt)
        o.damp_psi(dt, ch)
        if stage:
            o.average_saved()
        o.update_primitive()


def _linear_wave_error(orc, n, wave, rk2=False, axis=0):
    """L1 error of a small-amplitude wave along `axis` (0 x, 1 y, 2 z) after a quarter crossing.

    The reference advances with one forward-Euler stage per step (first order in time), so
    the CFL number is scaled with 1/n to expose the second-order spatial discretisation.
    rk2: the two-stage Runge-Kutta step at a fixed CFL number of 0.3, second order in time.
    The wave axis has n cells on [0, 1]; each transverse axis has 4 cells, 1.5/n and 0.75/n
    wide (the box is anisotropic, so a y or z spacing taken from the wrong axis changes the
    wave's speed and the error stops converging)."""
    widths = [1.0 / n] * 3
    widths[(axis + 1) % 3] = 1.5 / n
    widths[(axis + 2) % 3] = 0.75 / n
    ncell = [4, 4, 4]
    ncell[axis] = n
    o = orc.Oracle(nx=ncell[0], ny=ncell[1], nz=ncell[2], bc=(1, 1, 1, 1, 1, 1), xmin=(0, 0, 0),
                   xmax=tuple(w * c for w, c in zip(widths, ncell)),
                   gam=5.0 / 3.0, cfl=0.3 if rk2 else 0.1 * 32.0 / n, energy_flux=1, grid_shift=0,
                   stages=2 if rk2 else 1)
    o.generate_grid()
    gam = 5.0 / 3.0
    # p0 = 0.6/gam keeps c_s != c_A: at c_s = c_A the fast-speed discriminant of the reference
    # (cts^2 - 4 cs^2 Bn^2/rho, no guard) rounds below zero and its sqrt is NaN
    rho0, p0, bn = 1.0, 0.6 / gam, 1.0
    x = o.grid(axis + 1, True)
    amp = 1e-6
    P = o.P
    k = 2 * np.pi
    zero = 0 * x
    if wave == "sound":
        cs = np.sqrt(gam * p0 / rho0)
        bn = 0.0
        drho = amp * np.sin(k * x)
        rho = rho0 + drho
        vn = cs * drho / rho0
        p = p0 + cs * cs * drho
        speed, vt, bt = cs, zero, zero
    else:  # Alfven wave along B = bn e_axis: v_t = -dB_t / sqrt(rho)
        speed = bn / np.sqrt(rho0)
        rho, vn, p = rho0 + zero, zero, p0 + zero
        bt = amp * np.sin(k * x)
        vt = -bt / np.sqrt(rho0)
    # broadcast a profile along the wave axis over P[k, j, n, i]
    shape = [(1, 1, -1), (1, -1, 1), (-1, 1, 1)][axis]

    def put(nv, prof):
        P[:, :, nv, :] = np.broadcast_to(np.reshape(prof, shape), (o.ktot, o.jtot, o.itot))

    P[...] = 0.0
    t_ax = (axis + 1) % 3
    put(0, rho)
    put(1 + axis, vn)
    put(1 + t_ax, vt)
    put(5, p)
    put(4, p / ((gam - 1) * rho))
    put(6, np.sqrt(gam * p / rho))
    put(7 + axis, bn + zero)
    put(7 + t_ax, bt)
    put(11, 0.5 + zero)
    o.prim_to_cons_all()
    t_end = 0.25 / speed
    t = 0.0
    while t < t_end:
        dt = o.control_timestep()
        if t + dt > t_end:
            dt = t_end - t
        o.set_boundary()
        ch = o.evaluate_ch()
        if rk2:
            _rk2_step(o, dt, ch)
        else:
            for d in (1, 2, 3):
                o.numerical_flux(d, ch)
            o.update_conserv(dt)
            o.damp_psi(dt, ch)
            o.update_primitive()
        t += dt
    xs = x[2:2 + n] - speed * t_end
    line = [lambda v: o.P[2, 2, v, 2:2 + n], lambda v: o.P[2, 2:2 + n, v, 2], lambda v: o.P[2:2 + n, 2, v, 2]][axis]
    if wave == "sound":
        exact = rho0 + amp * np.sin(k * xs)
        got = line(0)
    else:
        exact = amp * np.sin(k * xs)
        got = line(7 + t_ax)
    return np.mean(np.abs(got - exact)) / amp


@pytest.mark.parametrize("wave", ["sound", "alfven"])
def test_second_order_convergence(orc, wave):
    e = [_linear_wave_error(orc, n, wave) for n in (32, 64, 128)]
    orders = np.log2(np.array(e[:-1]) / np.array(e[1:]))
    assert np.all(orders > 1.6), (e, orders)


@pytest.mark.parametrize("axis", [1, 2])
@pytest.mark.parametrize("wave", ["sound", "alfven"])
def test_second_order_convergence_y_and_z(orc, wave, axis):
    """the same waves along y and along z on anisotropic boxes: pins the y and z flux,
    reconstruction and update paths (their spacings and component rotations), which exact
    conservation cannot see"""
    e = [_linear_wave_error(orc, n, wave, axis=axis) for n in (32, 64, 128)]
    orders = np.log2(np.array(e[:-1]) / np.array(e[1:]))
    assert np.all(orders > 1.6), (e, orders)


@pytest.mark.parametrize("wave", ["sound", "alfven"])
def test_rk2_second_order_at_fixed_cfl(orc, wave):
    """the two-stage step converges at second order with the time step tied to the cell
    size (CFL 0.3); a dropped stage or average would leave first order in time"""
    e = [_linear_wave_error(orc, n, wave, rk2=True) for n in (32, 64, 128)]
    orders = np.log2(np.array(e[:-1]) / np.array(e[1:]))
    assert np.all(orders > 1.7), (e, orders)


def test_rk2_psi_damping_closed_form(orc):
    """a uniform state with uniform psi: every flux difference vanishes and only DampPsi acts,
    so each two-stage step multiplies psi by 1 - x + x^2/2 (x = dt 0.1 c_h / min(dx,dy,dz)),
    the second-order Taylor polynomial of exp(-x); one Euler step multiplies by 1 - x"""
    for stages in (1, 2):
        o = _box(orc, n=(8, 6, 4), xmin=(0, 0, 0), xmax=(1, 0.75, 1.0), gam=5.0 / 3.0, stages=stages, cfl=0.3,
                 bc=(1, 1, 1, 1, 1, 1))
        P = o.P
        rho, p, gam = 1.2, 0.9, 5.0 / 3.0
        P[:, :, 0] = rho
        P[:, :, 1:4] = np.array([0.3, -0.2, 0.1])[None, None, :, None]
        P[:, :, 5] = p
        P[:, :, 4] = p / ((gam - 1) * rho)
        P[:, :, 6] = np.sqrt(gam * p / rho)
        P[:, :, 7:10] = np.array([0.2, 0.1, -0.3])[None, None, :, None]
        P[:, :, 10] = 0.05
        P[:, :, 11] = 0.5
        o.prim_to_cons_all()
        dt, ch = o.control_timestep(), o.evaluate_ch()
        x = dt * 0.1 * ch / min(1 / 8, 0.75 / 6, 1 / 4)
        g = 1 - x + x * x / 2 if stages == 2 else 1 - x
        t, log = o.run(5)
        np.testing.assert_allclose(log, dt, rtol=1e-14)
        np.testing.assert_allclose(o.interior(o.P)[10], 0.05 * g ** 5, rtol=1e-13)
        np.testing.assert_allclose(o.interior(o.P)[0], rho, rtol=1e-14)


def test_rk2_run_is_the_two_stage_composition(orc):
    """orc_run with stages = 2 performs exactly the step the convergence test verifies"""
    kw = dict(n=(12, 10, 8), gam=5.0 / 3.0, cfl=0.3, energy_flux=1, stages=2)
    a, b = _box(orc, **kw), _box(orc, **kw)
    for o in (a, b):
        o.generate_problem_shear(bx0=0.1, a=0.08, amp=0.05)
    a.run(4)
    for _ in range(4):
        dt = b.control_timestep()
        b.set_boundary()
        _rk2_step(b, dt, b.evaluate_ch())
    np.testing.assert_array_equal(a.U, b.U)
    np.testing.assert_array_equal(a.interior(a.P), b.interior(b.P))


def test_prim_cons_roundtrip(orc):
    o = _box(orc, n=(6, 6, 6))
    o.generate_problem_shear(bx0=0.2)
    P0 = o.interior(o.P).copy()
    o.update_primitive()
    np.testing.assert_allclose(o.interior(o.P), P0, rtol=1e-13, atol=1e-15)


def test_splitmix64_known_values(orc):
    # published first outputs of splitmix64 seeded with 0 (state increments before mixing)
    assert orc.splitmix64(0) == 0xE220A8397B1DCDAF
    assert orc.splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_vissimple_t15_fixture_statistics():
    """the committed statistics of the reference's t = 15 sample (make_vissimple_t15_stats.py)
    hold the loose ranges read off those files: rho in [0.89, 1.09], p in [0.98, 1.28] (three
    digits), 22,500 cells of a 150^3 run's first interior z-plane, a well-mixed scalar"""
    import json
    with open(os.path.join(HERE, "golden", "vissimple_t15_stats.json")) as f:
        s = json.load(f)
    assert s["cells"] == 150 * 150
    assert s["d"]["min"] == pytest.approx(0.887) and s["d"]["max"] == pytest.approx(1.09)
    assert s["p"]["min"] == pytest.approx(0.978) and s["p"]["max"] == pytest.approx(1.28)
    for name in ("d", "vx", "vy", "p", "X1"):
        q = [s[name][k] for k in ("min", "q5", "q25", "q50", "q75", "q95", "max")]
        assert q == sorted(q) and s[name]["std"] > 0.0
    assert 0.2 < s["mix"] <= 0.25 and 0.0 < s["vy_rms"] < 0.5
