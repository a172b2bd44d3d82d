"""Independent textbook formulas used to pin the oracle (test-only, numpy).

Written from the published methods, not from the reference code:
  * ideal-MHD physical flux and fast magnetosonic speed (e.g. Toro, "Riemann Solvers and
    Numerical Methods for Fluid Dynamics", and Miyoshi & Kusano 2005 eq. 2-3);
  * HLLC for B = 0 (Toro, Spruce & Speares 1994; Toro 2009 sec. 10.4, eq. 10.39 star states);
  * HLLD with explicit wave-region branches (Miyoshi & Kusano 2005, J. Comput. Phys. 208:315,
    eqs. 38-51 and the flux selection of eq. 66).
Primitive vectors use the 12-entry order rho v1 v2 v3 e p cs B1 B2 B3 psi X; outputs are
Cartesian conserved-order flux vectors rho m1 m2 m3 E B1 B2 B3 psi rhoX.
"""
from __future__ import annotations

import numpy as np

RHO, V1, V2, V3, EI, PR, CS, B1, B2, B3, PSI, SC = range(12)


def conserved(q: np.ndarray) -> np.ndarray:
    rho = q[RHO]
    v = q[V1:V1 + 3]
    b = q[B1:B1 + 3]
    E = q[PR] / 1.0  # placeholder, overwritten below
    E = rho * q[EI] + 0.5 * rho * v @ v + 0.5 * b @ b
    return np.array([rho, *(rho * v), E, *b, q[PSI], rho * q[SC]])


def physical_flux(q: np.ndarray, d: int) -> np.ndarray:
    """F_d(U) of ideal MHD with the (v.B) B_n energy term; F(B_d) = F(psi) = 0."""
    rho = q[RHO]
    v = q[V1:V1 + 3]
    b = q[B1:B1 + 3]
    pt = q[PR] + 0.5 * b @ b
    E = rho * q[EI] + 0.5 * rho * v @ v + 0.5 * b @ b
    vn, bn = v[d], b[d]
    mom = rho * v * vn - b * bn
    mom[d] += pt
    ind = b * vn - v * bn
    ind[d] = 0.0
    return np.array([rho * vn, *mom, (E + pt) * vn - (v @ b) * bn, *ind, 0.0, rho * q[SC] * vn])


def fast_speed(q: np.ndarray, d: int) -> float:
    rho = q[RHO]
    b = q[B1:B1 + 3]
    a2 = q[CS] ** 2
    ca2 = (b @ b) / rho
    can2 = b[d] ** 2 / rho
    return float(np.sqrt(0.5 * (a2 + ca2 + np.sqrt((a2 + ca2) ** 2 - 4.0 * a2 * can2))))


def _rotate(q: np.ndarray, d: int):
    """(rho, u, v, w, p, Bx, By, Bz, E, sc) in the frame normal to d, transverse (d+1, d+2)."""
    t1, t2 = (d + 1) % 3, (d + 2) % 3
    rho = q[RHO]
    u, v, w = q[V1 + d], q[V1 + t1], q[V1 + t2]
    bx, by, bz = q[B1 + d], q[B1 + t1], q[B1 + t2]
    E = rho * q[EI] + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (bx * bx + by * by + bz * bz)
    return rho, u, v, w, q[PR], bx, by, bz, E, q[SC]


def _unrotate(fr: np.ndarray, d: int) -> np.ndarray:
    """rotated flux (rho, mu, mv, mw, E, Bu, Bv, Bw, sc) -> Cartesian 10-vector."""
    t1, t2 = (d + 1) % 3, (d + 2) % 3
    out = np.zeros(10)
    out[0] = fr[0]
    out[1 + d], out[1 + t1], out[1 + t2] = fr[1], fr[2], fr[3]
    out[4] = fr[4]
    out[5 + d], out[5 + t1], out[5 + t2] = fr[5], fr[6], fr[7]
    out[9] = fr[8]
    return out


def _rot_u(rho, u, v, w, E, bx, by, bz, sc):
    return np.array([rho, rho * u, rho * v, rho * w, E, bx, by, bz, rho * sc])


def _rot_f(rho, u, v, w, p, bx, by, bz, E, sc):
    pt = p + 0.5 * (bx * bx + by * by + bz * bz)
    vb = u * bx + v * by + w * bz
    return np.array([rho * u, rho * u * u + pt - bx * bx, rho * v * u - bx * by, rho * w * u - bx * bz,
                     (E + pt) * u - bx * vb, 0.0, by * u - bx * v, bz * u - bx * w, rho * sc * u])


def hlld(qL: np.ndarray, qR: np.ndarray, d: int) -> tuple[np.ndarray, str]:
    """Miyoshi & Kusano HLLD flux (branching form) and the region name that was used."""
    rl, ul, vl, wl, pl, bxl, byl, bzl, El, scl = _rotate(qL, d)
    rr, ur, vr, wr, pr, bxr, byr, bzr, Er, scr = _rotate(qR, d)
    bx = 0.5 * (bxl + bxr)
    cfl, cfr = fast_speed(qL, d), fast_speed(qR, d)
    SL = min(ul, ur) - max(cfl, cfr)
    SR = max(ul, ur) + max(cfl, cfr)
    ptl = pl + 0.5 * (bxl ** 2 + byl ** 2 + bzl ** 2)
    ptr = pr + 0.5 * (bxr ** 2 + byr ** 2 + bzr ** 2)
    UL = _rot_u(rl, ul, vl, wl, El, bxl, byl, bzl, scl)
    UR = _rot_u(rr, ur, vr, wr, Er, bxr, byr, bzr, scr)
    FL = _rot_f(rl, ul, vl, wl, pl, bxl, byl, bzl, El, scl)
    FR = _rot_f(rr, ur, vr, wr, pr, bxr, byr, bzr, Er, scr)
    FL[5] = FR[5] = 0.0
    if SL > 0:
        return _unrotate(FL, d), "L"
    if SR < 0:
        return _unrotate(FR, d), "R"
    # eq. 38 and 41
    den = (SR - ur) * rr - (SL - ul) * rl
    SM = ((SR - ur) * rr * ur - (SL - ul) * rl * ul - ptr + ptl) / den
    pts = ((SR - ur) * rr * ptl - (SL - ul) * rl * ptr + rl * rr * (SR - ur) * (SL - ul) * (ur - ul)) / den

    def star(r, u, v, w, by, bz, E, pt, S):
        rs = r * (S - u) / (S - SM)                                   # eq. 43
        dd = r * (S - u) * (S - SM) - bx * bx
        vs = v - bx * by * (SM - u) / dd                              # eq. 44
        ws = w - bx * bz * (SM - u) / dd                              # eq. 46
        bys = by * (r * (S - u) ** 2 - bx * bx) / dd                  # eq. 45
        bzs = bz * (r * (S - u) ** 2 - bx * bx) / dd                  # eq. 47
        Es = ((S - u) * E - pt * u + pts * SM + bx * (u * bx + v * by + w * bz - (SM * bx + vs * bys + ws * bzs))) / (S - SM)  # eq. 48
        return rs, vs, ws, bys, bzs, Es

    rsl, vsl, wsl, bysl, bzsl, Esl = star(rl, ul, vl, wl, byl, bzl, El, ptl, SL)
    rsr, vsr, wsr, bysr, bzsr, Esr = star(rr, ur, vr, wr, byr, bzr, Er, ptr, SR)
    sqL, sqR = np.sqrt(rsl), np.sqrt(rsr)
    SsL = SM - abs(bx) / sqL                                          # eq. 51
    SsR = SM + abs(bx) / sqR
    UsL = _rot_u(rsl, SM, vsl, wsl, Esl, bx, bysl, bzsl, scl)
    UsR = _rot_u(rsr, SM, vsr, wsr, Esr, bx, bysr, bzsr, scr)
    UsL[5] = UL[5]
    UsR[5] = UR[5]
    FsL = FL + SL * (UsL - UL)
    FsR = FR + SR * (UsR - UR)
    if SsL >= 0:
        return _unrotate(FsL, d), "L*"
    if SsR <= 0:
        return _unrotate(FsR, d), "R*"
    sg = np.sign(bx)
    inv = 1.0 / (sqL + sqR)
    vss = (sqL * vsl + sqR * vsr + (bysr - bysl) * sg) * inv          # eq. 59
    wss = (sqL * wsl + sqR * wsr + (bzsr - bzsl) * sg) * inv          # eq. 60
    byss = (sqL * bysr + sqR * bysl + sqL * sqR * (vsr - vsl) * sg) * inv  # eq. 61
    bzss = (sqL * bzsr + sqR * bzsl + sqL * sqR * (wsr - wsl) * sg) * inv  # eq. 62
    vbss = SM * bx + vss * byss + wss * bzss
    EssL = Esl - sqL * (SM * bx + vsl * bysl + wsl * bzsl - vbss) * sg  # eq. 63
    EssR = Esr + sqR * (SM * bx + vsr * bysr + wsr * bzsr - vbss) * sg
    UssL = _rot_u(rsl, SM, vss, wss, EssL, bx, byss, bzss, scl)
    UssR = _rot_u(rsr, SM, vss, wss, EssR, bx, byss, bzss, scr)
    UssL[5] = UL[5]
    UssR[5] = UR[5]
    if SM >= 0:
        return _unrotate(FsL + SsL * (UssL - UsL), d), "L**"
    return _unrotate(FsR + SsR * (UssR - UsR), d), "R**"


def hllc(qL: np.ndarray, qR: np.ndarray, d: int) -> np.ndarray:
    """Toro's HLLC for a B = 0 state with the same outer wave speeds as above."""
    rl, ul, vl, wl, pl, _, _, _, El, scl = _rotate(qL, d)
    rr, ur, vr, wr, pr, _, _, _, Er, scr = _rotate(qR, d)
    al, ar = qL[CS], qR[CS]
    SL = min(ul, ur) - max(al, ar)
    SR = max(ul, ur) + max(al, ar)
    Ss = (pr - pl + rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))  # Toro 10.37

    def ustar(r, u, v, w, p, E, sc, S):
        f = r * (S - u) / (S - Ss)
        return np.array([f, f * Ss, f * v, f * w, f * (E / r + (Ss - u) * (Ss + p / (r * (S - u)))), 0.0, 0.0, 0.0,
                         f * sc])                                                                    # Toro 10.39

    UL = _rot_u(rl, ul, vl, wl, El, 0.0, 0.0, 0.0, scl)
    UR = _rot_u(rr, ur, vr, wr, Er, 0.0, 0.0, 0.0, scr)
    FL = _rot_f(rl, ul, vl, wl, pl, 0.0, 0.0, 0.0, El, scl)
    FR = _rot_f(rr, ur, vr, wr, pr, 0.0, 0.0, 0.0, Er, scr)
    if SL >= 0:
        return _unrotate(FL, d)
    if SR <= 0:
        return _unrotate(FR, d)
    if Ss >= 0:
        return _unrotate(FL + SL * (ustar(rl, ul, vl, wl, pl, El, scl, SL) - UL), d)
    return _unrotate(FR + SR * (ustar(rr, ur, vr, wr, pr, Er, scr, SR) - UR), d)


def random_prim(rng: np.random.Generator, gam: float = 1.4, bscale: float = 1.0, vscale: float = 1.0) -> np.ndarray:
    """A physically consistent primitive vector (e = p/((gam-1)rho), cs = sqrt(gam p/rho))."""
    q = np.zeros(12)
    q[RHO] = rng.uniform(0.3, 3.0)
    q[V1:V1 + 3] = rng.uniform(-1.0, 1.0, 3) * vscale
    q[PR] = rng.uniform(0.2, 3.0)
    q[EI] = q[PR] / ((gam - 1.0) * q[RHO])
    q[CS] = np.sqrt(gam * q[PR] / q[RHO])
    q[B1:B1 + 3] = rng.uniform(-1.0, 1.0, 3) * bscale
    q[PSI] = rng.uniform(-0.1, 0.1)
    q[SC] = rng.uniform(0.0, 1.0)
    return q
