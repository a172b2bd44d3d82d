"""Benchmark: cell-updates/s of the MHD time step on one MI355X, with its HBM roofline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256 | --shape NX NY NZ]
                    [--mode 0|1] [--precision 0|1|2] [--stages 1|2 [--cfl C]] [--no-cpu]
                    [--no-serve] [--no-rk2] [--no-dropin]

Workload: a 256^3 FP64 Kelvin-Helmholtz shear-layer box (gamma 5/3, Bx = 0.1,
vx = 0.5 tanh(y/0.05), seeded vy kick; periodic x,z, reflecting y), advanced with the
reference's loop body (MUSCL + van Leer, HLLD, GLM, one forward-Euler stage, CFL
factor 0.05).  The state is generated on the device, so inputs are resident in HBM
before the timed region.  One step = one loop body over all 256^3 interior cells.
--stages 2 runs the two-stage Runge-Kutta step instead (CFL factor 0.3 unless --cfl is
given); one step is then both stages.

Printed JSON line: value = interior cells x steps / wall seconds of the timed region.  In
the timed region the step's kernels are launched one by one on the library's stream with a
HIP event after each launch (all recorded up front, one host wait at the end), so the
per-kernel durations come from the same steps as `value`.  graph_replay_ms_per_step times
the production loop (each step a replayed hipGraph, no events) for comparison.
roofline: algorithmic bytes of a cell update (read + write the 10 FP64 conserved
variables = 160 B) per launch of the dominant kernel / that kernel's average launch
duration in the timed region, against 8 TB/s.
valu: the same launch's VALU lane-instructions (rocprofv3 SQ_INSTS_VALU x 64, measured at
256^3) over its duration, against every SIMD issuing a wave64 instruction each 4 clocks at
2.4 GHz; the FP64 xy kernel is bound by this issue (two wavefronts per SIMD where a block
holds whole periodic x-rows, one elsewhere).
host_transfer: the reference's host-array crossings timed once (upload of P and U, download
of P, between pageable numpy arrays allocated once), and the rate if every step made that
round trip;
never part of `value`.
cpu_baseline: the oracle (C++/OpenMP restatement of the reference step) timed on this
host's cores on 16 steps (about 10 s) of the same 256^3 box: one thread per physical core
the process may use (at most OMP_NUM_THREADS), OMP_PROC_BIND=close, OMP_PLACES=cores, in a
child process run before the GPU is touched; the line names the CPU model and core counts.
ensemble / serving: the batched and request-serving measurements (skipped by --no-serve);
ensemble.rk2 repeats the ensemble with the two-stage step (skipped by --no-rk2 too).
dropin: the reference's own loop through the C++ drop-in layer at 256^3 in separate
processes (recorded onto the fused kernels, staged, AdvanceSteps; skipped by --no-dropin).
rk2: the same box advanced by the two-stage Runge-Kutta step at CFL factor 0.3 (the
configuration with "2-stage RK, CFL 0.3"), beside the reference's one-stage step; both
report simulated time per wall second (skipped by --no-rk2).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

BYTES_PER_CELL = 2 * 10 * 8      # read + write the FP64 conserved state (NV = 10)
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on MI355X, each
# counter scaled by what it reads for a copy kernel of known size on the same build
# (k_copy_interior, 1,342,177,280 B each way: fetch x2.0, write x1.0);
# summaries and commands in profiles/fused_fp64_256cube_pmc.md and
# profiles/fused_mixed_256cube_pmc.md.  Keyed by (n, mode, precision, kernel); a workload
# without a measurement reports null.
MEASURED_TRAFFIC = {
    (256, 0, 0, "fused_xy"): 694769.6 * 1024 * 2.0 + 1310721.7 * 1024 * 1.0,
    (256, 0, 0, "fused_z_update"): 1373216.0 * 1024 * 2.0 + 1310976.0 * 1024 * 1.0,
    (256, 0, 2, "fused_xy"): 719870.0 * 1024 * 2.0 + 655632.2 * 1024 * 1.0,
    (256, 0, 2, "fused_z_update"): 1112784.5 * 1024 * 2.0 + 1310976.0 * 1024 * 1.0,
}

# VALU lane-instructions per launch (SQ_INSTS_VALU x 64), measured with rocprofv3 --pmc on
# MI355X in the same profile runs; keyed like MEASURED_TRAFFIC.
MEASURED_VALU = {
    (256, 0, 0, "fused_xy"): 327.3e6 * 64,
    (256, 0, 0, "fused_z_update"): 209.7e6 * 64,
    (256, 0, 2, "fused_xy"): 286.7e6 * 64,
    (256, 0, 2, "fused_z_update"): 201.4e6 * 64,
}
# every SIMD issuing one wave64 VALU instruction per 4 clocks: 256 CUs x 4 SIMDs x 64 lanes
# / 4 clocks x 2.4 GHz
VALU_PEAK_LANE_INSTR_S = 256 * 4 * 64 / 4 * 2.4e9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_box(shape, mode_precision: int, stages: int = 1, cfl: float = 0.05):
    from mhdturbulence_amd import Box, Config
    nx, ny, nz = shape
    cfg = Config(nx=nx, ny=ny, nz=nz, bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5),
                 gam=5.0 / 3.0, cfl=cfl, energy_flux=1, grid_shift=0, precision=mode_precision, stages=stages)
    b = Box(cfg)
    b.generate_grid()
    b.generate_problem_shear(rho0=1.0, p0=1.0, v0=0.5, a=0.05, bx0=0.1, amp=1e-2, sigma=0.1, seed=20261203)
    return b


def host_cpu() -> dict:
    """CPU model, logical CPUs this process may run on, and the physical cores among them"""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        try:
            topo = f"/sys/devices/system/cpu/cpu{c}/topology/"
            with open(topo + "physical_package_id") as f1, open(topo + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        except OSError:
            cores.add(("cpu", c))
    return {"model": model, "logical_cpus": len(cpus), "physical_cores": len(cores)}


def _cpu_baseline_worker(shape, steps: int) -> dict:
    """oracle on this host's cores, same box, `steps` steps (after building its state); runs
    in the child process cpu_baseline starts with the OpenMP binding in its environment"""
    from oracle import oracle as orc
    nx, ny, nz = shape
    threads = int(os.environ["OMP_NUM_THREADS"])
    orc.lib().orc_set_threads(threads)
    o = orc.Oracle(nx=nx, ny=ny, nz=nz, bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5),
                   gam=5.0 / 3.0, cfl=0.05, energy_flux=1, grid_shift=0)
    o.generate_grid()
    o.generate_problem_shear(rho0=1.0, p0=1.0, v0=0.5, a=0.05, bx0=0.1, amp=1e-2, sigma=0.1, seed=20261203)
    o.run(1)  # first touch of every page outside the timed steps
    t0 = time.perf_counter()
    o.run(steps)
    wall = time.perf_counter() - t0
    return {"value": nx * ny * nz * steps / wall, "wall_s": wall}


def cpu_baseline(shape, steps: int) -> dict:
    """The oracle (C++/OpenMP restatement of the reference step) timed on this host: one
    thread per physical core this process may use (at most OMP_NUM_THREADS when that is
    set, the host's CPU share), OMP_PROC_BIND=close and OMP_PLACES=cores, in a child process
    so that the OpenMP runtime starts with that binding."""
    import subprocess
    cpu = host_cpu()
    threads = cpu["physical_cores"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit() and int(os.environ["OMP_NUM_THREADS"]) > 0:
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-worker", *map(str, shape),
                        str(steps)], env=env, capture_output=True, text=True, check=True)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    nx, ny, nz = shape
    name = f"{nx}^3" if nx == ny == nz else f"{nx}x{ny}x{nz}"
    return {"value": res["value"], "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the same {name} box (oracle C++/OpenMP restatement), {res['wall_s']:.1f} s",
            "cpu": cpu, "omp": {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}}


SHEAR_BASE = dict(bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5), gam=5.0 / 3.0, cfl=0.05,
                  energy_flux=1, grid_shift=0)


def serving_bench(count: int, seed: int = 20261203) -> dict:
    """Config 5's serving workload: shear-layer requests with n log-uniform in [32, 256]
    (multiples of 8), physical end times t_end uniform in [1, 6], per-request parameters as
    the ensemble's (a, Bx, amplitude, p from splitmix64(seed + id)), every solver taking the
    two-stage Runge-Kutta step at CFL factor 0.3.  `count` requests (cut from an open stream
    to fit the run; the line states it).

    saturated: every request queued at t = 0, served by the plain FP64 single-box solver one
    request at a time (the baseline), by the continuous-batching server in FP64 and in mixed
    precision (precision 2); rates and ratios are cell-updates per wall second.
    rate_4x: Poisson arrivals at 4 times the baseline's measured capacity (requests per
    second), through the baseline and the batched mixed server: latency under load."""
    from mhdturbulence_amd import Config
    from mhdturbulence_amd.serve import Server, poisson_stream, serve_one_at_a_time
    base = dict(SHEAR_BASE, stages=2, cfl=0.3)
    kw = dict(n_range=(32, 256), t_range=(1.0, 6.0), seed=seed)
    keys = ("cell_updates_per_s", "p50_latency_s", "p99_latency_s", "mean_latency_s", "wall_s", "failed")
    out = {"requests": count, "n_range": [32, 256], "t_end_range": [1.0, 6.0], "stages": 2, "cfl": 0.3,
           "note": f"{count} requests of an open stream, cut to fit the run"}
    reqs = poisson_stream(rate=0.0, count=count, **kw)
    out["sizes"] = [r.n for r in reqs]
    sat = {}
    log(f"[bench] serving: {count} requests one at a time (FP64 single-box solver)")
    sat["fp64_one_at_a_time"] = {k: v for k, v in serve_one_at_a_time(reqs, Config(**base, precision=0)).items()
                                 if k in keys}
    for prec, name in ((0, "batched_fp64"), (2, "batched_mixed")):
        log(f"[bench] serving: {count} requests, {name}")
        srv = Server(Config(**base, precision=prec), max_boxes=count, arena_bytes=48 << 30, chunk_steps=20)
        sat[name] = {k: v for k, v in srv.serve(reqs).items() if k in keys}
        srv.close()
    one = sat["fp64_one_at_a_time"]["cell_updates_per_s"]
    sat["batched_fp64_ratio"] = sat["batched_fp64"]["cell_updates_per_s"] / one
    sat["batched_mixed_ratio"] = sat["batched_mixed"]["cell_updates_per_s"] / one
    out["saturated"] = sat
    capacity = count / sat["fp64_one_at_a_time"]["wall_s"]
    rate = 4.0 * capacity
    row = {"arrival_rate_per_s": rate, "fp64_one_at_a_time_capacity_per_s": capacity}
    stream = poisson_stream(rate=rate, count=count, **kw)
    log(f"[bench] serving: Poisson arrivals at 4x capacity ({rate:.3f}/s)")
    row["fp64_one_at_a_time"] = {k: v for k, v in serve_one_at_a_time(stream, Config(**base, precision=0)).items()
                                 if k in keys}
    srv = Server(Config(**base, precision=2), max_boxes=count, arena_bytes=48 << 30, chunk_steps=20)
    row["batched_mixed"] = {k: v for k, v in srv.serve(stream).items() if k in keys}
    srv.close()
    row["p99_ratio"] = row["fp64_one_at_a_time"]["p99_latency_s"] / row["batched_mixed"]["p99_latency_s"]
    row["throughput_ratio"] = row["batched_mixed"]["cell_updates_per_s"] / row["fp64_one_at_a_time"][
        "cell_updates_per_s"]
    out["rate_4x"] = row
    return out


def ensemble_bench(count: int = 96, steps: int = 20, seed: int = 20261203, stages: int = 1) -> dict:
    """Batched ensemble: `count` boxes, box b drawn by ensemble_box(b) (each axis from
    {32, 48, 64, 96, 128}, a, Bx, amplitude, p from splitmix64(seed + b)), advanced together
    through the batched launches in FP64 and in mixed precision (precision 1, and 2 with the
    divergence stored in FP32), against the same boxes
    advanced one at a time by the single-box FP64 solver.  Rates are interior cells x steps
    over the wall time of the timed steps (boxes built and warmed up beforehand).
    stages = 2: every solver takes the two-stage Runge-Kutta step at CFL factor 0.3."""
    from mhdturbulence_amd import Batch, Box, Config
    from mhdturbulence_amd.serve import ENSEMBLE_SIZES, ensemble_box
    boxes = [ensemble_box(b, seed) for b in range(count)]
    cells = sum(sh[0] * sh[1] * sh[2] for sh, _ in boxes)
    base = dict(SHEAR_BASE, stages=stages, cfl=0.3 if stages == 2 else SHEAR_BASE["cfl"])
    out = {"boxes": count, "axis_sizes": list(ENSEMBLE_SIZES), "cells": cells, "steps": steps, "stages": stages}
    for prec, name in ((0, "batched_fp64"), (1, "batched_mixed"), (2, "batched_mixed_f32_divergence")):
        bt = Batch(Config(**base, precision=prec), max_boxes=count, arena_bytes=48 << 30)
        for sh, ic in boxes:
            bt.add_box(*sh, ic=ic)
        bt.step(2)
        t0 = time.perf_counter()
        gpu_ms = bt.step(steps)
        wall = time.perf_counter() - t0
        out[name] = {"cell_updates_per_s": cells * steps / wall, "gpu_ms_per_step": gpu_ms / steps}
        bt.close()
    wall = 0.0
    for sh, ic in boxes:
        box = Box(Config(nx=sh[0], ny=sh[1], nz=sh[2], **base))
        box.generate_grid()
        box.generate_problem_shear(**ic)
        box.run(2)
        t0 = time.perf_counter()
        box.run(steps)
        wall += time.perf_counter() - t0
        box.close()
    out["fp64_one_at_a_time"] = {"cell_updates_per_s": cells * steps / wall}
    out["batched_fp64_speedup"] = out["batched_fp64"]["cell_updates_per_s"] / out["fp64_one_at_a_time"][
        "cell_updates_per_s"]
    out["batched_mixed_speedup"] = out["batched_mixed"]["cell_updates_per_s"] / out["fp64_one_at_a_time"][
        "cell_updates_per_s"]
    out["batched_mixed_f32_divergence_speedup"] = out["batched_mixed_f32_divergence"]["cell_updates_per_s"] / out[
        "fp64_one_at_a_time"]["cell_updates_per_s"]
    return out


def dropin_bench(short: int = 20, long: int = 220) -> dict:
    """The reference's own loop (src_cpp_kokkos_host/main.cpp:228-267 call sequence, the
    reference KH problem with gamma 1.4 and B = 0) at 256^3 through the C++ drop-in layer
    (mhdturbulence_amd/mhd_kh_main_256, separate processes): its step functions recorded
    and run as the fused kernels (the default), each on its own staged kernels (--staged),
    and AdvanceSteps (mhd_run, the device loop).  Each variant runs `short` and `long`
    steps in two processes; ms per step is the difference of the driver's own loop timers
    over the difference of steps, so the one-time costs of a process (code loading, graph
    capture, first touch) cancel.  The loop's per-step host work (the dt read-back, printf,
    dt.log) stays in."""
    import re
    import subprocess
    import tempfile
    exe = os.path.join(HERE, "mhdturbulence_amd", "mhd_kh_main_256")
    out = {"workload": "reference KH problem 256^3 through the C++ drop-in", "steps": [short, long]}

    def loop_time(flags, steps, tmp):
        r = subprocess.run([exe, "--steps", str(steps), *flags], cwd=tmp, capture_output=True, text=True,
                           timeout=600, check=True)
        wall = float(re.search(r"sim time \[s\]: (\S+)", r.stdout).group(1))
        done = int(re.search(r"exiting main loop time=\S+, step=(\d+)", r.stdout).group(1)) + 1
        return wall, done

    with tempfile.TemporaryDirectory() as tmp:
        for name, flags in (("reference_loop_fused", []), ("reference_loop_staged", ["--staged"]),
                            ("advance_steps", ["--fused"])):
            w0, n0 = loop_time(flags, short, tmp)
            w1, n1 = loop_time(flags, long, tmp)
            ms = 1e3 * (w1 - w0) / (n1 - n0)
            out[name] = {"ms_per_step": ms, "cell_updates_per_s": 256 ** 3 / (ms * 1e-3),
                         "loop_s": [w0, w1], "steps_done": [n0, n1]}
    out["reference_loop_fused_over_advance_steps"] = (out["reference_loop_fused"]["ms_per_step"] /
                                                      out["advance_steps"]["ms_per_step"])
    return out


def rk2_bench(shape, precision: int, mode: int, euler_sim_rate: float, steps: int = 60) -> dict:
    """the two-stage Runge-Kutta step (CFL factor 0.3) on the same box: cell-updates/s (one
    update = both stages), per-kernel ms (both stages of a kernel summed) and simulated time
    per wall second against the one-stage step's"""
    cells = shape[0] * shape[1] * shape[2]
    box = make_box(shape, precision, 2, 0.3)
    t, _, _ = box.run(5, mode=mode)
    prof = box.profile_kernels(10, mode=mode)
    t, _, _ = box.run(2, time=t, mode=mode)
    t_before = t
    t0 = time.perf_counter()
    t, _, done = box.run(steps, time=t, mode=mode)
    wall = time.perf_counter() - t0
    out = {"stages": 2, "cfl": 0.3, "steps": done, "cell_updates_per_s": cells * done / wall,
           "ms_per_step": 1e3 * wall / max(done, 1), "kernels_ms": {k: round(v, 4) for k, v in prof.items()},
           "sim_time_per_wall_s": (t - t_before) / wall, "status": box.status()}
    out["sim_time_rate_vs_one_stage"] = out["sim_time_per_wall_s"] / euler_sim_rate
    box.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--mode", type=int, default=0, help="0 fused step kernels, 1 staged kernels")
    ap.add_argument("--precision", type=int, default=0, choices=(0, 1, 2),
                    help="0 FP64, 1 FP32 reconstruction + HLLD, 2 as 1 with the x-y divergence stored in FP32")
    ap.add_argument("--profile-steps", type=int, default=50,
                    help="steps of the graph-replay run timed after the measured steps")
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=12)
    ap.add_argument("--shape", type=int, nargs=3, metavar=("NX", "NY", "NZ"), default=None,
                    help="a non-cubic box instead of n^3 (e.g. one GPU's z-slab of a sharded box)")
    ap.add_argument("--stages", type=int, default=1, choices=(1, 2),
                    help="1: the reference's forward-Euler step; 2: two-stage Runge-Kutta")
    ap.add_argument("--cfl", type=float, default=None, help="CFL factor (default 0.05, or 0.3 with --stages 2)")
    ap.add_argument("--no-rk2", action="store_true")
    ap.add_argument("--no-dropin", action="store_true", help="skip the C++ drop-in driver runs")
    ap.add_argument("--cpu-baseline-worker", type=int, nargs=4, metavar=("NX", "NY", "NZ", "STEPS"),
                    help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_baseline_worker:
        nx, ny, nz, steps = args.cpu_baseline_worker
        print(json.dumps(_cpu_baseline_worker((nx, ny, nz), steps)), flush=True)
        return
    cfl = args.cfl if args.cfl is not None else (0.3 if args.stages == 2 else 0.05)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the whole box lives on one MI355X; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    n = args.n
    shape = tuple(args.shape) if args.shape else (n, n, n)
    cells = shape[0] * shape[1] * shape[2]
    cube = shape[0] == shape[1] == shape[2]
    shape_name = f"{n}^3" if cube and not args.shape else "x".join(map(str, shape))
    # the CPU baseline runs first, in a child process started before this process touches
    # the GPU
    cpu_line = None
    dropin = None
    if not args.no_dropin:
        log("[bench] the reference loop through the C++ drop-in at 256^3 (separate processes)")
        dropin = dropin_bench()
        log(f"[bench] drop-in: reference loop {dropin['reference_loop_fused']['ms_per_step']:.3f} ms/step fused, "
            f"{dropin['reference_loop_staged']['ms_per_step']:.3f} staged, "
            f"{dropin['advance_steps']['ms_per_step']:.3f} AdvanceSteps")
    if not args.no_cpu:
        log(f"[bench] cpu baseline: oracle, {args.cpu_steps} steps of {shape_name}")
        cpu_line = cpu_baseline(shape, args.cpu_steps)
        log(f"[bench] cpu baseline {cpu_line['value']:.4e} cell-updates/s on {cpu_line['cores']} cores")
    log(f"[bench] building {shape_name} box on the device")
    box = make_box(shape, args.precision, args.stages, cfl)
    log(f"[bench] device bytes {box.device_bytes() / 2**30:.2f} GiB; warmup {args.warmup} steps")
    t = 0.0
    t, _, _ = box.run(args.warmup, time=t, mode=args.mode)
    log(f"[bench] timing {args.steps} steps (an event after every launch, one host wait at the end)")
    t0 = time.perf_counter()
    prof = box.profile_kernels(args.steps, mode=args.mode)  # returns after the stream drained
    wall = time.perf_counter() - t0
    status = box.status()
    value = cells * args.steps / wall
    ms_per_step = 1e3 * wall / args.steps
    log(f"[bench] {value:.4e} cell-updates/s, {ms_per_step:.3f} ms/step (kernels {sum(prof.values()):.3f} ms/step)")
    # the production loop (each step a replayed graph, no events) over --profile-steps steps
    t_before = t
    t0 = time.perf_counter()
    t, _, done = box.run(args.profile_steps, time=t, mode=args.mode)
    graph_wall = time.perf_counter() - t0
    graph_ms = 1e3 * graph_wall / max(done, 1)
    sim_rate = (t - t_before) / graph_wall
    log(f"[bench] graph replay {graph_ms:.3f} ms/step, {sim_rate:.4f} simulated time units per wall second")
    # host buffers across the boundary (not part of `value`): the reference's FieldArray
    # h2d of P and U and the d2h of P, timed once each after a warm-up transfer, into host
    # arrays allocated once and already touched (as the reference's P and U are)
    P_host, U_host = box.download_prim(), box.download_cons()
    t0 = time.perf_counter()
    box.upload_prim(P_host)
    box.upload_cons(U_host)
    up_ms = 1e3 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    box.download_prim(out=P_host)
    down_ms = 1e3 * (time.perf_counter() - t0)
    del P_host, U_host
    pcie = {"upload_P_U_ms": round(up_ms, 2), "download_P_ms": round(down_ms, 2),
            "cell_updates_per_s_if_each_step_round_trips": cells / ((ms_per_step + up_ms + down_ms) * 1e-3)}
    log(f"[bench] host transfers: upload P+U {up_ms:.1f} ms, download P {down_ms:.1f} ms")
    dominant = max(prof, key=prof.get)
    # per-kernel times sum both stages of a two-stage step; the roofline is per launch
    dom_ms = prof[dominant] / (1 if dominant in ("step_end", "reduce_dt_ch") else args.stages)
    step_ms = sum(prof.values()) / args.stages
    # algorithmic bytes of one launch of the dominant kernel: the per-cell-update figure (160 B,
    # read + write of the FP64 conserved state) times the cells that launch processes (every
    # interior cell); "step" relates the same bytes to the whole step's kernel time (per
    # stage for a two-stage step)
    alg_bytes = BYTES_PER_CELL * cells
    achieved = alg_bytes / (dom_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": (MEASURED_TRAFFIC.get((shape[0], args.mode, args.precision, dominant))
                            if cube and args.stages == 1 else None),
                "kernel": dominant,
                "kernel_ms": round(dom_ms, 4),
                "step": {"achieved": round(alg_bytes / (step_ms * 1e-3) / 1e9, 2),
                         "frac": round(alg_bytes / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    # the secondary ceiling: VALU issue of the same kernel launch
    valu_n = MEASURED_VALU.get((shape[0], args.mode, args.precision, dominant)) if cube and args.stages == 1 else None
    valu = None
    if valu_n:
        rate = valu_n / (dom_ms * 1e-3)
        valu = {"kernel": dominant, "lane_instr_per_launch": valu_n, "achieved": rate,
                "peak": VALU_PEAK_LANE_INSTR_S, "unit": "lane-instr/s", "frac": round(rate / VALU_PEAK_LANE_INSTR_S, 4)}
    out = {
        "metric": "cell-updates/s (% HBM roofline) at 256^3 FP64; p99 latency of batched requests",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {0: "f64", 1: "f64 state, f32 reconstruction+HLLD",
                  2: "f64 state, f32 reconstruction+HLLD, f32 stored divergence"}[args.precision],
        "data": "synthetic (device-generated seeded shear-layer state)",
        "config": {"workload": f"KH shear layer {shape_name}, {'fused' if args.mode == 0 else 'staged'} step kernels",
                   "cells": cells, "gamma": 5.0 / 3.0, "bx0": 0.1, "cfl": cfl, "stages": args.stages,
                   "energy_flux": "(v.B)B_n", "parallelism": "single GPU"},
        "roofline": roofline,
        "valu": valu,
        "kernels_ms": {k: round(v, 4) for k, v in prof.items()},
        "graph_replay_ms_per_step": round(graph_ms, 4),
        "sim_time_per_wall_s": sim_rate,
        "host_transfer": pcie,
        "status": status,
        "sim_time": t,
    }
    del box
    if not args.no_rk2 and args.stages == 1:
        log(f"[bench] two-stage Runge-Kutta step at CFL 0.3, {shape_name}")
        out["rk2"] = rk2_bench(shape, args.precision, args.mode, sim_rate)
    if not args.no_serve:
        log("[bench] batched ensemble: 96 boxes, batched FP64 / mixed vs one at a time")
        out["ensemble"] = ensemble_bench()
        if not args.no_rk2:
            log("[bench] the same ensemble with the two-stage Runge-Kutta step")
            out["ensemble"]["rk2"] = ensemble_bench(stages=2)
        log(f"[bench] serving: {args.serve_requests} requests, batched vs one at a time")
        out["serving"] = serving_bench(args.serve_requests)
        r4 = out["serving"]["rate_4x"]
        out["p99_latency_s"] = {"batched_mixed_at_4x_capacity": r4["batched_mixed"]["p99_latency_s"],
                                "fp64_one_at_a_time_at_4x_capacity": r4["fp64_one_at_a_time"]["p99_latency_s"]}
    out["dropin"] = dropin
    out["cpu_baseline"] = cpu_line
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
