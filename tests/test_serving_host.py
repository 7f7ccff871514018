"""Host-side scheduling of the serving loop (no device): admission by
footprint, bounded skip-ahead, error propagation, statistics."""

from __future__ import annotations

import threading
import time

import numpy as np
import pytest

from realspace_exatdiagonalization_py_amd.serving import SectorSolveServer, SolveRequest, poisson_arrivals, serve_stream


class _Probe:
    def __init__(self):
        self.lock = threading.Lock()
        self.in_use = 0
        self.peak = 0
        self.running = 0
        self.peak_running = 0
        self.order = []

    def solve(self, server, req, ctx):
        size, dur, fail = req.tag
        with self.lock:
            self.in_use += size
            self.running += 1
            self.peak = max(self.peak, self.in_use)
            self.peak_running = max(self.peak_running, self.running)
            self.order.append(req.sector)
        time.sleep(dur)
        with self.lock:
            self.in_use -= size
            self.running -= 1
        if fail:
            raise ValueError(f"sector {req.sector} failed")
        return np.array([float(req.sector)]), None, 1, None


def _server(probe, workers=4, budget=100, max_skip=8, policy="fifo"):
    return SectorSolveServer(workers, hbm_budget=budget, catalog_cache_bytes=0, max_skip=max_skip, policy=policy,
                             solve_fn=probe.solve, context_factory=lambda: None,
                             estimate_fn=lambda r: r.tag[0])


def _req(i, size, dur=0.01, fail=False):
    return SolveRequest(model=None, n_filled=0, sector=i, tag=(size, dur, fail))


def test_admission_respects_budget():
    probe = _Probe()
    with _server(probe, workers=6, budget=100) as srv:
        futs = [srv.submit(_req(i, 30, 0.02)) for i in range(24)]
        res = [f.result(timeout=30) for f in futs]
    assert [r.eigenvalues[0] for r in res] == list(range(24))
    assert probe.peak <= 100          # at most three 30-byte solves at once
    assert probe.peak_running == 3
    st = srv.stats()
    assert st["completed"] == 24 and st["peak_admitted_bytes"] <= 100
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_oversized_request_runs_alone():
    probe = _Probe()
    with _server(probe, workers=4, budget=100) as srv:
        futs = [srv.submit(_req(0, 500, 0.03))] + [srv.submit(_req(i, 10, 0.01)) for i in range(1, 6)]
        for f in futs:
            f.result(timeout=30)
    assert probe.peak_running >= 1
    assert probe.order[0] == 0


def test_skip_ahead_is_bounded():
    """A large head request lets at most max_skip smaller ones pass it."""
    probe = _Probe()
    with _server(probe, workers=2, budget=100, max_skip=2) as srv:
        first = srv.submit(_req(0, 60, 0.05))        # occupies 60 of 100
        time.sleep(0.01)
        big = srv.submit(_req(1, 60, 0.01))          # does not fit while 0 runs
        small = [srv.submit(_req(i, 10, 0.05)) for i in range(2, 8)]
        for f in [first, big] + small:
            f.result(timeout=30)
    pos = probe.order.index(1)
    passed = [s for s in probe.order[:pos] if s >= 2]
    assert len(passed) <= 2


def test_errors_reach_the_caller_and_the_loop_continues():
    probe = _Probe()
    with _server(probe, workers=2) as srv:
        bad = srv.submit(_req(0, 10, 0.01, fail=True))
        good = srv.submit(_req(1, 10, 0.01))
        with pytest.raises(ValueError, match="sector 0 failed"):
            bad.result(timeout=30)
        assert good.result(timeout=30).eigenvalues[0] == 1.0


def test_poisson_stream_and_shutdown():
    probe = _Probe()
    srv = _server(probe, workers=3)
    arr = poisson_arrivals(10, 500.0, seed=1)
    assert np.all(np.diff(arr) > 0)
    res = serve_stream(srv, [_req(i, 10, 0.005) for i in range(10)], arr)
    assert len(res) == 10 and all(r.latency >= r.service >= 0 for r in res)
    srv.shutdown()
    with pytest.raises(RuntimeError):
        srv.submit(_req(99, 1))


def test_sjf_serves_the_smallest_admissible_first():
    probe = _Probe()
    with _server(probe, workers=1, budget=1000, policy="sjf") as srv:
        blocker = srv.submit(_req(0, 5, 0.05))  # queued requests wait behind it
        time.sleep(0.01)
        futs = [srv.submit(_req(i, size, 0.001)) for i, size in ((1, 50), (2, 10), (3, 30), (4, 20))]
        for f in [blocker] + futs:
            f.result(timeout=30)
    assert probe.order == [0, 2, 4, 3, 1]


def test_sjf_aging_bounds_the_wait():
    """A large request is passed over by smaller ones at most max_skip times."""
    probe = _Probe()
    with _server(probe, workers=1, budget=1000, max_skip=2, policy="sjf") as srv:
        blocker = srv.submit(_req(0, 5, 0.05))
        time.sleep(0.01)
        big = srv.submit(_req(1, 90, 0.001))
        small = [srv.submit(_req(i, 10, 0.001)) for i in range(2, 8)]
        for f in [blocker, big] + small:
            f.result(timeout=30)
    assert probe.order.index(1) <= 3  # blocker, then at most two small ones


def test_unknown_policy_is_rejected():
    with pytest.raises(ValueError):
        SectorSolveServer(1, hbm_budget=10, policy="lifo", solve_fn=lambda *a: None, context_factory=lambda: None)


def test_twisted_term_lists_are_cached_per_model_object():
    from realspace_exatdiagonalization_py_amd import models

    m1 = models.build_spinful_hubbard_second_quantized_model([2, 2], 1.0, 4.0)
    m2 = models.build_spinful_hubbard_second_quantized_model([2, 2], 1.0, 4.0)
    srv = _server(_Probe(), workers=1)
    try:
        r = lambda m, tw: SolveRequest(model=m, n_filled=2, sector=(0, 0), twist=tw)  # noqa: E731
        a = srv._bilinear_terms(r(m1, (0.25, 0.0)))
        assert srv._bilinear_terms(r(m1, (0.25, 0.0))) is a
        b = srv._bilinear_terms(r(m2, (0.25, 0.0)))
        assert b is not a and b == a  # another model object gets its own entry
        assert srv._bilinear_terms(r(m1, (0.0, 0.0))) is m1.bilinear_terms  # no twist: the model's own terms
        srv._max_term_sets = 2
        c = srv._bilinear_terms(r(m1, (0.5, 0.0)))
        assert len(srv._terms) == 2 and c != a
        assert srv._bilinear_terms(r(m1, (0.25, 0.0))) is not a  # the oldest entry was dropped
    finally:
        srv.shutdown()
