"""Long Kelvin-Helmholtz run with the reference's t-prof diagnostics.

The reference KH problem (src_cpp_kokkos_host/main.cpp:66-184: double shear layer, gamma 1.4,
B = 0, sinusoidal vy seed of amplitude 1e-3) on an n^3 box, advanced by the fused step to
t_end.  Every `interval` of simulated time the RealTimeAnalysis row of
src_cpp_omp_device/main.cpp:23-69 is appended to t-prof.csv in the reference's format:
time, mix, sqrt(<vy^2>), and the reference growth curve A exp(Gamma t) with A = 1.2e-3,
Gamma = 1.49 (src_cpp_omp_device/main.cpp:54-55).

The growth rate of sqrt(<vy^2>) is fitted (least squares of its logarithm) over
[fit0, fit1], and the steepest such rate over any one-time-unit window is reported too;
both are compared with Gamma.

    python -m mhdturbulence_amd.tprof [--n 150] [--t-end 15] [--interval 0.1] [--out DIR]
                                      [--stages 1|2] [--cfl C] [--gamma G] [--grid-shift 0|1]
                                      [--slice FILE]

--stages 2 advances with the two-stage Runge-Kutta step instead of the reference's
forward-Euler step (CFL factor 0.3 unless --cfl is given).  --gamma and --grid-shift set the
ratio of specific heats and the face formula (the Fortran variants take 5/3 and the
unshifted faces).  --slice writes the first interior z-plane of the final state in the
column format of the reference's analysis/vissimple samples: x y d vx vy p phi X1
(phi = psi), three significant digits, after a header of time and dt.

Prints one JSON object (wall time, steps, seconds per cell per step, fitted rate).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from .solver import Box, Config

A_REF, GAMMA_REF = 1.2e-3, 1.49


def write_slice(path: str, box: Box, t: float, dt: float):
    """the first interior z-plane of the box's primitives as the analysis/vissimple samples
    hold theirs (one row per cell, i fastest: x y d vx vy p phi X1)"""
    P = box.interior(box.download_prim())  # (nv, nz, ny, nx)
    x1b = box.get_grid()[1][2:2 + box.cfg.nx]
    x2b = box.get_grid()[3][2:2 + box.cfg.ny]
    with open(path, "w") as f:
        f.write(f"# {t:12.3E} {dt:12.3E}\n# {box.cfg.nx}\n# {box.cfg.ny}\n")
        f.write("# x y d vx vy p phi X1\n")
        for j in range(box.cfg.ny):
            for i in range(box.cfg.nx):
                q = P[:, 0, j, i]
                f.write(" ".join(f"{v:12.3E}" for v in (x1b[i], x2b[j], q[0], q[1], q[2], q[5], q[10], q[11])) + "\n")


def run(n: int, t_end: float, interval: float, out_dir: str | None, fit=(0.5, 3.0), precision: int = 0,
        log=None, stages: int = 1, cfl: float | None = None, gam: float = 1.4, grid_shift: int = 1,
        slice_path: str | None = None) -> dict:
    if cfl is None:
        cfl = 0.3 if stages == 2 else 0.05
    # the KH configuration
    box = Box(Config(nx=n, ny=n, nz=n, precision=precision, stages=stages, cfl=cfl, gam=gam, grid_shift=grid_shift))
    box.generate_grid()
    box.generate_problem_kh()
    rows = []
    t = 0.0
    d = box.diagnostics()
    rows.append((t, d[1], d[2]))
    steps = 0
    wall = 0.0
    # one sample after each chunk of steps; the chunk length follows dt so that samples
    # come about `interval` apart
    t1, dts, done = box.run(1, time=t)
    log_dt = dts
    steps += done
    per = max(1, round(interval / dts[0]))
    t = t1
    while t < t_end:
        t0 = time.perf_counter()
        t_new, log_dt, done = box.run(per, time=t, tmax=t_end)
        wall += time.perf_counter() - t0
        steps += done
        t = t_new
        d = box.diagnostics()
        rows.append((t, d[1], d[2]))
        if log and len(rows) % 10 == 0:
            log(f"  t={t:.3f} steps={steps} sqrt<vy^2>={d[2]:.4e} ({wall:.1f} s)")
        if done < per:
            break
        per = max(1, round(interval / log_dt[-1]))
    status = box.status()
    if slice_path:
        write_slice(slice_path, box, t, float(log_dt[-1]) if len(log_dt) else 0.0)
    box.close()
    arr = np.array(rows)
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        with open(os.path.join(out_dir, "t-prof.csv"), "w") as f:
            f.write(f"# Gamma= {GAMMA_REF:24.16E}\n# 1:time 2:mix 3:v_y 4:exp(2Gamma*t)\n")
            for tt, mix, vy in rows:
                f.write(f" {tt:24.16E} {mix:24.16E} {vy:24.16E} {A_REF * np.exp(GAMMA_REF * tt):24.16E}\n")
    sel = (arr[:, 0] >= fit[0]) & (arr[:, 0] <= fit[1])
    gamma = float(np.polyfit(arr[sel, 0], np.log(arr[sel, 2]), 1)[0]) if sel.sum() >= 3 else float("nan")
    # the steepest growth over any window of one time unit (the linear phase's rate)
    peak, t_peak = float("nan"), float("nan")
    for t_a in arr[:, 0]:
        w = (arr[:, 0] >= t_a) & (arr[:, 0] <= t_a + 1.0)
        if w.sum() >= 5 and t_a + 1.0 <= arr[-1, 0] + 1e-12:
            g = float(np.polyfit(arr[w, 0], np.log(arr[w, 2]), 1)[0])
            if not peak >= g:
                peak, t_peak = g, float(t_a)
    return {"n": n, "t_end": t_end, "stages": stages, "cfl": cfl, "gamma": gam, "grid_shift": grid_shift,
            "time_reached": t, "steps": steps,
            "status": status, "wall_s": wall,
            "s_per_cell_step": wall / (n ** 3 * max(steps, 1)), "fit_window": list(fit), "growth_rate": gamma,
            "peak_growth_rate": peak, "peak_window_start": t_peak, "reference_growth_rate": GAMMA_REF,
            "vy_rms_peak": float(arr[:, 2].max()), "t_of_peak": float(arr[int(arr[:, 2].argmax()), 0]),
            "samples": len(rows)}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=150)
    ap.add_argument("--t-end", type=float, default=15.0)
    ap.add_argument("--interval", type=float, default=0.1)
    ap.add_argument("--fit", type=float, nargs=2, default=[0.5, 3.0])
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--out", default=None, help="directory for t-prof.csv")
    ap.add_argument("--stages", type=int, default=1, choices=(1, 2))
    ap.add_argument("--cfl", type=float, default=None)
    ap.add_argument("--gamma", type=float, default=1.4)
    ap.add_argument("--grid-shift", type=int, default=1, choices=(0, 1))
    ap.add_argument("--slice", default=None, help="file for the final first interior z-plane")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    print(json.dumps(run(a.n, a.t_end, a.interval, a.out, tuple(a.fit), a.precision, log, a.stages, a.cfl,
                         a.gamma, a.grid_shift, a.slice)))


if __name__ == "__main__":
    main()
