"""Accuracy of the mixed-precision step (FP32 reconstruction + HLLD, FP64 state, update,
recovery and reductions) against the FP64 step on the KH shear layer.

Two measurements per Bx in {0.05, 0.1, 0.2} (domain [-0.5, 0.5]^3, rho = p = 1,
vx = 0.5 tanh(y/0.05), seeded vy kick, gamma 5/3, CFL factor 0.05):

* fields: an n^3 box (default 512^3) advanced `steps` steps in each precision from the same
  device-generated state; relative L1 differences of rho and p, both max |divB|, and the
  simulated times reached;
* growth: a growth_n^3 box (default 128^3) advanced to t_end in each precision, with
  sqrt(<vy^2>) sampled every `sample_dt`; the rate is the least-squares slope of
  ln sqrt(<vy^2>) over [fit0, fit1] (the seeded cell-scale kick first decays, the shear
  mode grows out of it later).

    python -m mhdturbulence_amd.accuracy [--n 512 --steps 100] [--growth-n 128 --t-end 3]
                                         [--stages 2 [--cfl 0.3]]

--stages 2 runs both precisions with the two-stage Runge-Kutta step (CFL factor 0.3 unless
--cfl is given) instead of the reference's forward-Euler step.  --mixed 2 compares the mixed
mode that also stores the x-y flux divergence in FP32 (precision 2) instead of precision 1.

Prints one JSON object.  The FP64 path is the one whose results equal the oracle's (the
parity tests); this tool measures how far the mixed path departs from it.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .solver import Box, Config

BASE = dict(bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5), gam=5.0 / 3.0, cfl=0.05,
            energy_flux=1, grid_shift=0, stages=1)
RHO, PRE = 0, 5  # primitive slots of density and pressure


def _box(n: int, precision: int, bx0: float, **step) -> Box:
    b = Box(Config(nx=n, ny=n, nz=n, precision=precision, **{**BASE, **step}))
    b.generate_grid()
    b.generate_problem_shear(rho0=1.0, p0=1.0, v0=0.5, a=0.05, bx0=bx0, amp=1e-2, sigma=0.1, seed=20261203)
    return b


def field_errors(n: int, steps: int, bx0: float, log=None, mixed: int = 1, **step) -> dict:
    res = {}
    prims = {}
    for prec in (0, mixed):
        b = _box(n, prec, bx0, **step)
        t0 = time.perf_counter()
        t, dts, done = b.run(steps)
        wall = time.perf_counter() - t0
        d = b.diagnostics()
        P = b.download_prim()  # (ktot, jtot, 12, itot)
        prims[min(prec, 1)] = P[2:b.ktot - 3, 2:b.jtot - 3, [RHO, PRE], 2:b.itot - 3]  # interior (nz, ny, 2, nx)
        del P
        res["fp64" if prec == 0 else "mixed"] = {"time": t, "steps": done, "max_divB": float(d[3]),
                                                 "vy_rms": float(d[2]), "wall_s": wall}
        b.close()
        if log:
            log(f"  n={n} Bx={bx0} prec={prec}: t={t:.6g} divB={d[3]:.3e} ({wall:.1f} s)")
    a, m = prims[0], prims[1]
    for k, name in ((0, "rho"), (1, "p")):
        res[f"rel_l1_{name}"] = float(np.abs(m[:, :, k] - a[:, :, k]).sum() / np.abs(a[:, :, k]).sum())
        res[f"rel_max_{name}"] = float((np.abs(m[:, :, k] - a[:, :, k]) / np.abs(a[:, :, k])).max())
    return res


def growth(n: int, t_end: float, sample_dt: float, bx0: float, fit=(1.5, 3.0), log=None, mixed: int = 1,
           **step) -> dict:
    res = {}
    for prec in (0, mixed):
        b = _box(n, prec, bx0, **step)
        t = 0.0
        ts, vy = [0.0], [float(b.diagnostics()[2])]
        # steps per sample from the first step's dt (dt changes slowly in the linear phase)
        t1, dts, _ = b.run(1, time=t)
        per = max(1, int(sample_dt / max(dts[0], 1e-12)))
        t = t1
        while t < t_end:
            t, _, done = b.run(per, time=t, tmax=t_end)
            ts.append(t)
            vy.append(float(b.diagnostics()[2]))
            if done < per:
                break
        ts, vy = np.array(ts), np.array(vy)
        sel = (ts >= fit[0]) & (ts <= fit[1])
        gamma = float(np.polyfit(ts[sel], np.log(vy[sel]), 1)[0]) if sel.sum() >= 3 else float("nan")
        res["fp64" if prec == 0 else "mixed"] = {"growth_rate": gamma, "t": ts.tolist(), "vy_rms": vy.tolist()}
        b.close()
        if log:
            log(f"  growth n={n} Bx={bx0} prec={prec}: Gamma={gamma:.4f} over t in {fit}")
    g0, g1 = res["fp64"]["growth_rate"], res["mixed"]["growth_rate"]
    res["growth_rate_rel_diff"] = abs(g1 - g0) / abs(g0) if g0 == g0 and g0 != 0 else float("nan")
    vy0, vy1 = np.array(res["fp64"]["vy_rms"]), np.array(res["mixed"]["vy_rms"])
    k = min(len(vy0), len(vy1))
    res["vy_rms_max_rel_diff"] = float(np.max(np.abs(vy1[:k] - vy0[:k]) / np.abs(vy0[:k])))
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--growth-n", type=int, default=128)
    ap.add_argument("--t-end", type=float, default=3.0)
    ap.add_argument("--sample-dt", type=float, default=0.05)
    ap.add_argument("--fit", type=float, nargs=2, default=[1.5, 3.0])
    ap.add_argument("--bx", type=float, nargs="+", default=[0.05, 0.1, 0.2])
    ap.add_argument("--stages", type=int, default=1, choices=(1, 2))
    ap.add_argument("--mixed", type=int, default=1, choices=(1, 2), help="the mixed precision level compared")
    ap.add_argument("--cfl", type=float, default=None)
    a = ap.parse_args(argv)
    step = {"stages": a.stages, "cfl": a.cfl if a.cfl is not None else (0.3 if a.stages == 2 else 0.05)}

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    out = {"step": step, "mixed_precision": a.mixed, "fields": {"n": a.n, "steps": a.steps},
           "growth": {"n": a.growth_n, "t_end": a.t_end, "fit": a.fit}}
    for bx in a.bx:
        out["fields"][f"Bx={bx:g}"] = field_errors(a.n, a.steps, bx, log, mixed=a.mixed, **step)
        g = growth(a.growth_n, a.t_end, a.sample_dt, bx, fit=tuple(a.fit), log=log, mixed=a.mixed, **step)
        row = {k: v for k, v in g.items() if k not in ("fp64", "mixed")}
        row.update({f"{p}_growth_rate": g[p]["growth_rate"] for p in ("fp64", "mixed")})
        row["t"] = [round(x, 4) for x in g["fp64"]["t"]]
        row["fp64_vy_rms"] = [float(f"{x:.6e}") for x in g["fp64"]["vy_rms"]]
        out["growth"][f"Bx={bx:g}"] = row
    print(json.dumps(out))


if __name__ == "__main__":
    main()
