"""Serving scheduler on the CPU with a mock executor: per-stream order, the
slot limit, HBM-budget admission, rejection of streams that can never fit,
and continuous admission of requests that arrive while others run."""
import threading
import time

import pytest

from dcvc_amd.serving import Request, Server, poisson_trace


class MockExecutor:
    def __init__(self, footprints, work_s=0.01):
        self.fp = footprints
        self.work_s = work_s
        self.lock = threading.Lock()
        self.active_streams = set()
        self.order = {}
        self.violations = []

    def footprint(self, stream):
        return self.fp[stream]

    def run(self, r):
        with self.lock:
            if r.stream in self.active_streams:
                self.violations.append(("overlap", r.stream))
            self.active_streams.add(r.stream)
            self.order.setdefault(r.stream, []).append(r.rid)
        time.sleep(self.work_s)
        with self.lock:
            self.active_streams.discard(r.stream)
        return r.rid


def test_order_slots_and_completion():
    ex = MockExecutor({s: 1 for s in range(8)})
    reqs = poisson_trace(rate=400, duration=0.5, n_streams=8, seed=1)
    srv = Server(ex, slots=3, hbm_budget=100)
    srv.run(reqs)
    assert all(r.t_done >= r.t_start >= r.t_arrive - 1e-6 for r in reqs)
    assert srv.max_running <= 3 and not ex.violations
    for s, rids in ex.order.items():            # each stream served in arrival order
        assert rids == sorted(rids)
    st = Server.stats(reqs)
    assert st["served"] == len(reqs) and st["errors"] == 0


def test_budget_limits_resident_streams_and_refuses_what_cannot_fit():
    fp = {0: 40, 1: 40, 2: 40, 3: 500}
    ex = MockExecutor(fp, work_s=0.005)
    reqs = [Request(i, 0.0, i % 3, "encode", 0) for i in range(9)] + [Request(9, 0.0, 3, "decode", 0)]
    srv = Server(ex, slots=4, hbm_budget=100)
    srv.run(reqs, realtime=False)
    assert srv.max_resident_bytes <= 100
    # stream 3 can never fit; stream 2 arrives after 0 and 1 hold 80 of 100 bytes
    assert reqs[9].error and all(r.error for r in reqs if r.stream == 2)
    assert all(not r.error and r.result == r.rid for r in reqs if r.stream in (0, 1))
    assert Server.stats(reqs)["served"] == 6


def test_requests_arriving_mid_run_overlap_running_ones():
    ex = MockExecutor({s: 1 for s in range(4)}, work_s=0.2)
    reqs = [Request(0, 0.0, 0, "encode", 0), Request(1, 0.05, 1, "encode", 0)]
    srv = Server(ex, slots=2, hbm_budget=10)
    srv.run(reqs)
    # the second request started while the first was still running
    assert reqs[1].t_start < reqs[0].t_done
    assert srv.max_running == 2


def test_stats_split_first_requests_from_warm_ones():
    """Each (stream, kind)'s first request counts as a first request; the
    warm percentiles cover the rest."""
    from dcvc_amd.serving import Request, Server
    reqs = []
    t = 0.0
    for i, (stream, kind, lat) in enumerate([(0, "encode", 0.100), (1, "decode", 0.080), (0, "encode", 0.005),
                                             (0, "decode", 0.090), (1, "decode", 0.004), (0, "encode", 0.006)]):
        reqs.append(Request(rid=i, t_arrive=t, stream=stream, kind=kind, qp=0, t_start=t, t_done=t + lat))
        t += 0.01
    st = Server.stats(reqs)
    assert st["first_requests"] == 3
    assert abs(st["first_request_mean_ms"] - 90.0) < 1e-6
    assert st["warm_p99_ms"] < 6.1 and st["warm_p50_ms"] == 5.0
    assert st["p99_ms"] > 90
