"""Serving on the GPU: the request queue with continuous batching (bwpm_queue_* through
`bwpm_amd.serve`). Every result in the mix must equal the same solve run alone bit for bit
(FP64 and FP32, interface and Born solves, more requests than lane slots, so requests join
running lanes). The queue must report its latency statistics, admit by HBM footprint and report
failures per request."""

from __future__ import annotations

import numpy as np
import pytest

from bwpm_amd import _core, serve, workloads

pytestmark = pytest.mark.gpu


def _alone(p: workloads.Problem, precision: str) -> dict:
    s = _core.GpuSolver(*p.solver_args(), fluctuation=p.fluctuation, precision=precision)
    try:
        return s.run(p.field, p.iterations)
    finally:
        s.close()


def _mix() -> list[tuple[workloads.Problem, str]]:
    reqs = []
    for i, (n, nz, its) in enumerate([(64, 40, 1), (64, 32, 2), (64, 24, 3), (64, 40, 2), (64, 16, 1),
                                      (64, 28, 2)]):
        reqs.append((workloads.tilted_3d(n=n, nz=nz, iterations=its, polarisation="xy"[i % 2]), "fp64"))
    reqs.append((serve.planar_3d(96, 30, 2, 1.0, 1.0, serve.SILVER, layer=(10, 13)), "fp64"))
    reqs.append((serve.planar_3d(96, 20, 1, 1.3, 1.0, 1.5), "fp64"))
    reqs.append((workloads.plain_2d(ny=256, nz=50, iterations=2), "fp64"))
    reqs.append((workloads.born_3d(n=32, nz=12, iterations=1, sigma=0.01, seed=5), "fp64"))
    reqs.append((workloads.tilted_3d(n=64, nz=30, iterations=2), "fp32"))
    reqs.append((workloads.tilted_3d(n=64, nz=26, iterations=2, geometry=True), "mixed"))
    reqs.append((workloads.tilted_3d(n=64, nz=34, iterations=1, polarisation="y"), "mixed"))
    reqs.append((workloads.tilted_3d(n=64, nz=20, iterations=0), "fp64"))
    return reqs


def _same(a: dict, b: dict) -> None:
    assert np.array_equal(a["transmitted"], b["transmitted"])
    assert np.array_equal(a["reflected"], b["reflected"])
    assert len(a["transmitted_iter"]) == len(b["transmitted_iter"])
    for x, y in zip(a["transmitted_iter"] + a["reflected_iter"], b["transmitted_iter"] + b["reflected_iter"]):
        assert np.array_equal(x, y)
    assert a["stored_values"] == b["stored_values"]


def test_results_in_the_mix_equal_the_solves_alone() -> None:
    reqs = _mix()
    # 3 slots per lane and 8 slab steps per chunk: the 64² lane has 6 FP64 requests, so later ones
    # join the running lane between chunks
    with serve.Server(slots=3, chunk=8) as srv:
        tickets = [srv.submit(p, precision=prec) for p, prec in reqs]
        results = [srv.result(t, timeout=120) for t in tickets]
        st = srv.stats()
    for (p, prec), res in zip(reqs, results):
        print("compare", p.name, prec, flush=True)
        _same(res, _alone(p, prec))
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    assert st["updates_per_s"] > 0
    assert st["used_bytes"] == 0 and st["running"] == 0 and st["pending"] == 0
    assert 0 < st["peak_bytes"] <= st["budget_bytes"]


def test_admission_by_hbm_footprint() -> None:
    small = [workloads.tilted_3d(n=64, nz=24, iterations=1, polarisation="xy"[i % 2]) for i in range(4)]
    big = workloads.tilted_3d(n=256, nz=64, iterations=2)
    budget = 8e6  # a 64² solve fits, two at once do not, the 256² solve never does
    with serve.Server(hbm_budget=budget, slots=4, chunk=16) as srv:
        tickets = [srv.submit(p) for p in small]
        t_big = srv.submit(big)
        for p, t in zip(small, tickets):
            _same(srv.result(t, timeout=120), _alone(p, "fp64"))
        with pytest.raises(RuntimeError, match="HBM budget"):
            srv.result(t_big, timeout=120)
        st = srv.stats()
    assert st["completed"] == 4 and st["failed"] == 1
    assert st["peak_bytes"] <= budget


def test_request_errors_are_per_ticket() -> None:
    good = workloads.tilted_3d(n=64, nz=16, iterations=1)
    bad = workloads.tilted_3d(n=64, nz=16, iterations=1)
    bad.labels = bad.labels.copy()
    bad.labels[3, 5, 7] = 9  # no such index
    with serve.Server(slots=2, chunk=8) as srv:
        tb = srv.submit(bad)
        tg = srv.submit(good)
        with pytest.raises(ValueError):
            srv.result(tb, timeout=120)
        _same(srv.result(tg, timeout=120), _alone(good, "fp64"))
        with pytest.raises(ValueError, match="unknown ticket"):
            srv.state(tb)
        with pytest.raises(ValueError, match="field"):
            srv._q.submit(good.ny, good.nx, good.dy, good.dx, good.dz, good.k0, good.indices, good.labels,
                          good.index_before, None, good.field[:10], 1)


def test_poisson_stream_of_the_served_mix() -> None:
    reqs = [serve.served_request(100 + i, max_n=128, max_nz=240, born_n=32) for i in range(16)]
    for p in reqs:  # a 2D request of 4096 points keeps the stream short
        if p.nx == 1 and p.ny > 1024:
            p.labels = p.labels[:, :1024]
            p.field = p.field[:1024]
            p.ny = 1024
    seen = {}
    with serve.Server(slots=8, chunk=32) as srv:
        rep = serve.run_stream(srv, reqs, rate_hz=40.0, seed=7,
                               on_result=lambda t, p, r: seen.setdefault(p.name + str(t), r))
    assert rep.submitted == len(reqs) and not rep.failures
    assert rep.stats["completed"] == len(reqs)
    assert rep.stats["p50_ms"] <= rep.stats["p99_ms"]
    assert len(seen) == len(reqs)
    for r in seen.values():
        assert np.all(np.isfinite(r["transmitted"]))


def test_replayed_lane_graphs_match_eager_solves() -> None:
    # slab steps in a homogeneous medium make uniform chunks: the lane captures one graph and
    # replays it for every chunk of every request, with eager work of other requests in between
    # (deep enough that chunks remain after the tile tuner has timed its candidates)
    reqs = []
    for i in range(6):
        p = workloads.tilted_3d(n=64, nz=200 + 8 * i, iterations=1 + i % 2, polarisation="xy"[i % 2])
        p.labels = np.zeros_like(p.labels)
        p.indices = [1.33 + 0j]
        p.index_before = 1.33 + 0j
        reqs.append(p)
    other = workloads.tilted_3d(n=96, nz=40, iterations=1)
    before = _core.lane_graph_stats()
    with serve.Server(slots=2, chunk=16) as srv:
        tickets = [srv.submit(p) for p in reqs]
        t_other = srv.submit(other)
        results = [srv.result(t, timeout=120) for t in tickets]
        _same(srv.result(t_other, timeout=120), _alone(other, "fp64"))
    for p, res in zip(reqs, results):
        _same(res, _alone(p, "fp64"))
    after = _core.lane_graph_stats()
    assert after["built"] > before["built"] and after["replays"] > before["replays"] + 6
    assert after["fallbacks"] == before["fallbacks"]


def test_few_generic_solves_in_a_wide_lane() -> None:
    # a 96² grid runs the generic kernels: with 2 of 16 slots busy the lane launches the narrow
    # geometry, with 16 busy the wide one; every result equals its solve alone
    few = [workloads.tilted_3d(n=96, nz=24 + 4 * i, iterations=1 + i, polarisation="xy"[i]) for i in range(2)]
    hom = serve.planar_3d(96, 30, 1, 1.0, 1.0, 1.0)
    many = [workloads.tilted_3d(n=96, nz=12 + i, iterations=1) for i in range(16)]
    with serve.Server(slots=16, chunk=8) as srv:
        t_few = [srv.submit(p) for p in few + [hom]]
        r_few = [srv.result(t, timeout=120) for t in t_few]
        t_many = [srv.submit(p) for p in many]
        r_many = [srv.result(t, timeout=120) for t in t_many]
    for p, r in zip(few + [hom] + many, r_few + r_many):
        _same(r, _alone(p, "fp64"))
