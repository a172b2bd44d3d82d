"""Request server: continuous batching of simulation requests on one MI355X.

Requests (a shear-layer box of n^3 cells run to time t_end) arrive over time.  The server

* admits a pending request as soon as a box slot and an arena region of its HBM
  footprint are free (FIFO, no reordering);
* keeps every admitted box resident in HBM and advances all live boxes together, one
  batched launch per kernel per step (mhd_batch_*);
* retires a box when it passes its t_end (each box stops on its own clock), returns its
  t-prof diagnostics (volume, mix, sqrt(<vy^2>), max|divB|) and frees its region for the
  next request;
* drops a request its client cancels (Request.cancel_at): still pending, it is never
  admitted; live, its box is removed at the next chunk boundary and its slot and region
  go to the next request.  The other boxes' results do not change.

`serve()` replays a request stream against the wall clock and reports per-request latency
(completion - arrival), p50 / p99 and aggregate cell-updates/s.  `serve_one_at_a_time()` is
the baseline: the plain single-box solver taking the same stream one request at a time.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .solver import Batch, Config


@dataclass
class Request:
    rid: int
    n: int
    t_end: float
    ic: dict
    arrival: float = 0.0
    # seconds after the stream starts at which the client cancels the request (None: never)
    cancel_at: float | None = None


@dataclass
class Result:
    rid: int
    n: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    sim_time: float
    diagnostics: list = field(default_factory=list)
    # mhd_batch_box_state status: 0 = clean; bit 0 = the box met a cell with rho <= 0 or a
    # non-finite CFL bound and stopped there, before its t_end
    status: int = 0
    prim: np.ndarray | None = None  # final primitive state (reference layout), if kept
    cancelled: bool = False  # dropped at its cancel_at, before t_end (state as it stood)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


def request_params(rid: int, seed: int = 20261203) -> dict:
    """per-request shear-layer parameters: a in [0.02,0.1], Bx in [0,0.2], amplitude in
    [1e-3,1e-2], p in [0.5,5], drawn from splitmix64(seed + rid)"""
    s = seed + rid
    u = []
    for _ in range(4):
        s = splitmix64(s)
        u.append((s >> 11) * 2.0 ** -53)
    return dict(a=0.02 + 0.08 * u[0], bx0=0.2 * u[1], amp=1e-3 + 9e-3 * u[2], p0=0.5 + 4.5 * u[3],
                seed=(seed + rid) & 0xFFFFFFFFFFFFFFFF)


ENSEMBLE_SIZES = (32, 48, 64, 96, 128)


def ensemble_box(b: int, seed: int = 20261203) -> tuple[tuple[int, int, int], dict]:
    """Box b of the batched ensemble: (nx, ny, nz) each drawn from ENSEMBLE_SIZES and the
    shear-layer parameters (a, Bx, amplitude, p) from the splitmix64 sequence started at
    seed + b: a in [0.02,0.1], Bx in [0,0.2], amplitude in [1e-3,1e-2], p in [0.5,5]"""
    s = seed + b
    u = []
    for _ in range(7):
        s = splitmix64(s)
        u.append((s >> 11) * 2.0 ** -53)
    shape = tuple(ENSEMBLE_SIZES[min(int(x * len(ENSEMBLE_SIZES)), len(ENSEMBLE_SIZES) - 1)] for x in u[:3])
    ic = dict(a=0.02 + 0.08 * u[3], bx0=0.2 * u[4], amp=1e-3 + 9e-3 * u[5], p0=0.5 + 4.5 * u[6],
              seed=(seed + b) & 0xFFFFFFFFFFFFFFFF)
    return shape, ic


def poisson_stream(rate: float, count: int, seed: int = 20261203, n_range=(32, 256), t_range=(1.0, 6.0),
                   multiple: int = 8) -> list[Request]:
    """Poisson arrivals at `rate` per second (rate 0: every request at t = 0); n log-uniform
    in n_range rounded to a multiple of `multiple`; t_end uniform in t_range.  Sizes and end
    times come from one generator and arrival gaps from another, so a seed gives the same
    requests at every rate."""
    rng = np.random.default_rng(seed)
    gaps = np.random.default_rng([seed, 1])
    t = 0.0
    out = []
    for rid in range(count):
        t += gaps.exponential(1.0 / rate) if rate > 0 else 0.0
        n = math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))
        n = int(max(n_range[0], min(n_range[1], round(n / multiple) * multiple)))
        out.append(Request(rid, n, float(rng.uniform(*t_range)), request_params(rid, seed), t))
    return out


def _summary(results: list, wall: float, cell_updates: int, gpu_s: float) -> dict:
    """latency percentiles over the completed requests; cancelled ones are counted apart"""
    done = [r for r in results if not r.cancelled]
    lat = np.array([r.latency for r in done] or [0.0])
    return {
        "requests": len(done),
        "cancelled": len(results) - len(done),
        "wall_s": wall,
        "gpu_s": gpu_s,
        "cell_updates": cell_updates,
        "cell_updates_per_s": cell_updates / wall,
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()),
        "failed": sum(1 for r in results if r.status),
        "results": results,
    }


def serve_one_at_a_time(requests: list[Request], cfg: Config, chunk_steps: int = 500, clock=time.perf_counter,
                        keep_prim: bool = False) -> dict:
    """The plain single-box solver serving the stream one request at a time, in arrival
    order: each request gets its own Box (mhd_create, the device problem generator, mhd_run
    in chunks of chunk_steps with tmax = t_end, diagnostics), as a user of the
    one-box-per-job reference would run it; the next request starts when it is done."""
    from .solver import Box
    reqs = sorted(requests, key=lambda r: r.arrival)
    t0 = clock()
    results: list[Result] = []
    cell_updates = 0
    gpu_ms = 0.0
    for r in reqs:
        now = clock() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        adm = clock() - t0
        c = Config(**{**cfg.__dict__, "nx": r.n, "ny": r.n, "nz": r.n})
        box = Box(c)
        box.generate_grid()
        box.generate_problem_shear(**r.ic)
        t, steps = 0.0, 0
        while True:
            t, _, done = box.run(chunk_steps, time=t, tmax=r.t_end)
            gpu_ms += box.last_run_timing()[0]
            steps += done
            if done < chunk_steps or t > r.t_end:
                break
        diag = box.diagnostics()
        prim = box.download_prim() if keep_prim else None
        status = box.status()
        box.close()
        cell_updates += steps * r.n ** 3
        results.append(Result(r.rid, r.n, r.arrival, adm, clock() - t0, steps, t, diag.tolist(), status, prim))
    return _summary(results, clock() - t0, cell_updates, gpu_ms / 1e3)


class Server:
    def __init__(self, cfg: Config, max_boxes: int = 64, arena_bytes: int = 64 << 30, chunk_steps: int = 10,
                 tilings: dict | None = None):
        self.batch = Batch(cfg, max_boxes, arena_bytes)
        self.chunk = chunk_steps
        for (nx, ny, nz), (sw, ly, lz) in (tilings or {}).items():
            self.batch.set_tiling(nx, ny, nz, sw, ly, lz)

    def close(self):
        self.batch.close()

    def footprint(self, n: int) -> int:
        return self.batch.box_bytes(n, n, n)

    def serve(self, requests: list[Request], clock=time.perf_counter, log=None, keep_prim: bool = False) -> dict:
        """keep_prim: each Result also carries its final primitive state (downloaded as the box
        retires; for small boxes, as in tests)"""
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = clock()
        nxt = 0
        pending: list[Request] = []
        live: dict[int, tuple[Request, float]] = {}
        results: list[Result] = []
        cell_updates = 0
        gpu_ms = 0.0
        def cancelled(r: Request, now: float) -> bool:
            return r.cancel_at is not None and now >= r.cancel_at

        while len(results) < len(reqs):
            now = clock() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                pending.append(reqs[nxt])
                nxt += 1
            # cancellations: pending requests are dropped, live boxes removed with their state
            for r in [r for r in pending if cancelled(r, now)]:
                pending.remove(r)
                results.append(Result(r.rid, r.n, r.arrival, now, now, 0, 0.0, cancelled=True))
            for bid in [b for b, (r, _) in live.items() if cancelled(r, now)]:
                r, adm = live.pop(bid)
                st = self.batch.box_state(bid)
                diag, prim = self.batch.finish_box(bid, want_prim=keep_prim)
                self.batch.remove_box(bid)
                cell_updates += st["steps"] * r.n ** 3
                results.append(Result(r.rid, r.n, r.arrival, adm, clock() - t0, st["steps"], st["time"],
                                      diag.tolist(), st["status"], prim, cancelled=True))
                if log:
                    log(f"cancelled rid={r.rid} n={r.n} after {st['steps']} steps")
            # admission in arrival order while a slot and a region of the footprint are free
            while pending:
                r = pending[0]
                bid = self.batch.try_add_box(r.n, r.n, r.n, tmax=r.t_end, ic=r.ic)
                if bid is None:
                    if not live:
                        raise RuntimeError(f"request {r.rid} (n={r.n}) does not fit the arena even alone")
                    break
                live[bid] = (r, clock() - t0)
                pending.pop(0)
            if not live:
                if nxt < len(reqs) and not pending:
                    time.sleep(max(0.0, reqs[nxt].arrival - (clock() - t0)))
                continue
            gpu_ms += self.batch.step(self.chunk)
            for bid in self.batch.poll():
                r, adm = live.pop(bid)
                st = self.batch.box_state(bid)
                diag, prim = self.batch.finish_box(bid, want_prim=keep_prim)
                self.batch.remove_box(bid)
                cell_updates += st["steps"] * r.n ** 3
                res = Result(r.rid, r.n, r.arrival, adm, clock() - t0, st["steps"], st["time"], diag.tolist(),
                             st["status"], prim)
                results.append(res)
                if log:
                    log(f"done rid={r.rid} n={r.n} steps={st['steps']} latency={res.latency:.3f}s")
        return _summary(results, clock() - t0, cell_updates, gpu_ms / 1e3)
