"""Continuous-batching server: every request completes, requests overlap on the device, and
each result is the same alone and in the mix."""
import numpy as np
import pytest

from mhdturbulence_amd import Config
from mhdturbulence_amd.serve import Request, Server, poisson_stream, request_params

pytestmark = pytest.mark.gpu

BASE = dict(bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5), gam=5.0 / 3.0, energy_flux=1,
            grid_shift=0)


def test_stream_completes_and_overlaps():
    reqs = poisson_stream(rate=200.0, count=10, seed=5, n_range=(16, 48), t_range=(0.002, 0.006))
    srv = Server(Config(**BASE), max_boxes=6, arena_bytes=1 << 30, chunk_steps=5)
    out = srv.serve(reqs)
    srv.close()
    assert out["requests"] == 10
    res = {r.rid: r for r in out["results"]}
    for r in reqs:
        x = res[r.rid]
        assert x.finished >= x.admitted >= r.arrival - 1e-9
        assert x.sim_time > r.t_end and x.steps > 0
        assert np.isfinite(x.diagnostics).all()
    # some requests were admitted while others were still running
    spans = sorted((x.admitted, x.finished) for x in out["results"])
    assert any(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1))
    assert out["p99_latency_s"] >= out["p50_latency_s"] > 0


def test_result_same_alone_and_in_mix():
    target = Request(99, 24, 0.004, request_params(99, 7), 0.0)
    others = [Request(k, n, 0.006, request_params(k, 7), 0.0) for k, n in enumerate((32, 16, 40))]
    alone = Server(Config(**BASE), max_boxes=4, arena_bytes=1 << 30, chunk_steps=3)
    ra = alone.serve([target])["results"][0]
    alone.close()
    mixed = Server(Config(**BASE), max_boxes=4, arena_bytes=1 << 30, chunk_steps=3)
    out = mixed.serve([others[0], target, others[1], others[2]])
    mixed.close()
    rm = [r for r in out["results"] if r.rid == 99][0]
    assert rm.steps == ra.steps and rm.sim_time == ra.sim_time
    assert rm.diagnostics == ra.diagnostics  # bitwise: per-box arithmetic only


def test_admission_waits_for_memory():
    # an arena that holds one 32^3 box at a time: the second request waits for the first
    probe = Server(Config(**BASE), max_boxes=4, arena_bytes=64 << 20, chunk_steps=4)
    need = probe.footprint(32)
    probe.close()
    srv = Server(Config(**BASE), max_boxes=4, arena_bytes=int(need * 1.5), chunk_steps=4)
    reqs = [Request(0, 32, 0.003, request_params(0), 0.0), Request(1, 32, 0.003, request_params(1), 0.0)]
    out = srv.serve(reqs)
    srv.close()
    a, b = sorted(out["results"], key=lambda r: r.admitted)
    assert b.admitted >= a.finished - 1e-6


@pytest.fixture(scope="module")
def orc(oracle_lib):
    return oracle_lib


def _oracle_request(orc, r, stages=2, cfl=0.3, steps=None):
    """an oracle run of request r: the same box and scheme, stepped until the step that takes
    time past t_end (src_cpp_kokkos_host/main.cpp:265), or exactly `steps` steps"""
    o = orc.Oracle(nx=r.n, ny=r.n, nz=r.n, **BASE, stages=stages, cfl=cfl)
    o.generate_grid()
    o.generate_problem_shear(**r.ic)
    t, n = 0.0, 0
    while (steps is None and t <= r.t_end) or (steps is not None and n < steps):
        t, _ = o.run(1, t)
        n += 1
    return o, t, n


@pytest.mark.parametrize("server", ["batched_fp64", "batched_mixed", "one_at_a_time"])
def test_served_requests_match_oracle(orc, server):
    """config 5's scheme (two-stage step at CFL 0.3) on small requests served together:
    each request's final state against an oracle run of that request alone, FP64 at
    rel L1 <= 1e-10 with the same step count and end time, precision 2 at the mixed bar"""
    import parity
    from mhdturbulence_amd.serve import serve_one_at_a_time
    reqs = [Request(k, n, t_end, request_params(k, 11), 0.0)
            for k, (n, t_end) in enumerate(((16, 0.09), (24, 0.05), (16, 0.13), (32, 0.04)))]
    cfg = Config(**BASE, stages=2, cfl=0.3, precision=2 if server == "batched_mixed" else 0)
    if server == "one_at_a_time":
        out = serve_one_at_a_time(reqs, cfg, chunk_steps=5, keep_prim=True)
    else:
        srv = Server(cfg, max_boxes=4, arena_bytes=1 << 30, chunk_steps=7)
        out = srv.serve(reqs, keep_prim=True)
        srv.close()
    assert out["requests"] == len(reqs) and out["failed"] == 0
    for res in out["results"]:
        r = reqs[res.rid]
        if server == "batched_mixed":
            o, t, n = _oracle_request(orc, r, steps=res.steps)
        else:
            o, t, n = _oracle_request(orc, r)
            assert res.steps == n and res.sim_time == pytest.approx(t, rel=1e-12)
        assert res.sim_time > r.t_end and res.status == 0
        err = parity.rel_l1(o.interior(res.prim), o.interior(o.P), "prim")
        if server == "batched_mixed":
            assert err[0] <= 1e-6 and err[5] <= 1e-6 and np.all(err <= 1e-4), err
        else:
            assert err.max() <= 1e-10, (r, err)


def test_cancelled_requests_leave_the_others_unchanged():
    """a live request cancelled after its first chunk of steps and a pending one cancelled
    before admission: both are returned as cancelled, their slot and region are freed, and
    the other requests end with the same bits as when served without them.  The clock is a
    counter (one unit per reading), so the cancellations fall at fixed points of the loop."""
    def counter():
        n = [0]

        def clock():
            n[0] += 1
            return float(n[0])
        return clock

    keep = [Request(0, 16, 0.006, request_params(0, 9), 0.0), Request(1, 24, 0.005, request_params(1, 9), 0.0)]
    live_cancel = Request(2, 32, 10.0, request_params(2, 9), 0.0, cancel_at=4.5)
    pending_cancel = Request(3, 16, 10.0, request_params(3, 9), 0.0, cancel_at=4.5)
    srv = Server(Config(**BASE), max_boxes=3, arena_bytes=1 << 30, chunk_steps=3)
    out = srv.serve(keep + [live_cancel, pending_cancel], clock=counter())
    avail, total = srv.batch.free_bytes()
    srv.close()
    assert avail == total
    assert out["requests"] == 2 and out["cancelled"] == 2
    res = {r.rid: r for r in out["results"]}
    assert res[2].cancelled and res[2].steps == 3 and res[2].sim_time < live_cancel.t_end
    assert res[3].cancelled and res[3].steps == 0
    ref_srv = Server(Config(**BASE), max_boxes=3, arena_bytes=1 << 30, chunk_steps=3)
    ref = {r.rid: r for r in ref_srv.serve(keep, clock=counter())["results"]}
    ref_srv.close()
    for k in (0, 1):
        assert not res[k].cancelled and res[k].sim_time > keep[k].t_end
        assert res[k].steps == ref[k].steps and res[k].sim_time == ref[k].sim_time
        assert res[k].diagnostics == ref[k].diagnostics
