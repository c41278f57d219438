"""The serving loop: requests on resident hypergraphs, admitted by HBM
footprint and run concurrently by worker threads, give the oracle's
refinement of the request, and the same part vector and metrics as the same
request run alone through a fresh context."""
import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd.engine import SCHED_BAL, SCHED_REF, Engine, Server, schedule
from oracle import oracle as O

import oracle_refine as R

pytestmark = pytest.mark.gpu

OBJ = ("TV", "CN", "TM", "MSM", "MSV", "MSRV")


def _sched(metric, hg, K, C, max_pw, min_pw, cap):
    if metric == "CN":
        return None
    if metric == "TV":
        return schedule(SCHED_BAL, 4096, C, cap, 0, cap)
    return schedule(SCHED_REF, hg.nov if metric == "TM" else 1024, C, max_pw, min_pw)


def _workload():
    graphs = [H.colnet(3000, H.DEG_UNIFORM, 2, 9, seed=11), H.workload("c2", scale=0.01),
              H.colnet(2000, H.DEG_UNIFORM, 1, 3, seed=3), H.colnet(5000, H.DEG_FIXED, 6, seed=4)]
    reqs = []
    for i in range(18):
        g = i % len(graphs)
        hg = graphs[g]
        metric = OBJ[i % 6]
        K = (16, 64, 32)[i % 3]
        C, max_pw, min_pw = H.ref_constants(hg.nov, K)
        part = H.initial_partition_for(metric, hg, K)   # the objective's own initial stream
        reqs.append((g, metric, K, part, _sched(metric, hg, K, C, max_pw, min_pw, H.balance_cap(hg.nov, K, 0.05)), 3))
    return graphs, reqs


def _alone(hg, metric, K, part, sched, passes):
    with Engine(0) as e:
        e.upload(hg)
        e.set_partition(K, part)
        e.refine(metric, passes, sched)
        return e.get_state().part, e.metrics()


def _serve(graphs, reqs, budget, workers):
    with Server(0, hbm_budget=budget, workers=workers) as srv:
        for hg in graphs:
            srv.add_graph(hg)
        print(f"serving {len(reqs)} requests on {len(graphs)} hypergraphs, budget {budget}, {workers} workers",
              flush=True)
        rids = [srv.submit(g, metric, K, part, sched, passes) for g, metric, K, part, sched, passes in reqs]
        out = [srv.wait(r) for r in rids]
        return out, srv.stats()


def test_serve_same_alone_and_in_the_mix():
    graphs, reqs = _workload()
    out, stats = _serve(graphs, reqs, 1 << 36, 6)
    assert stats["completed"] == len(reqs)
    assert stats["peak_inflight"] > 1   # requests did overlap
    for (info, part), (g, metric, K, p0, sched, passes) in zip(out, reqs):
        want_part, want_m = _alone(graphs[g], metric, K, p0, sched, passes)
        assert np.array_equal(part, want_part), (g, metric, K)
        assert info["metrics"] == want_m, (g, metric, K)
        assert info["start_ms"] >= info["submit_ms"] and info["end_ms"] >= info["start_ms"]
        # the oracle's refinement of the request's state (the first-appearance
        # lists of its part vector)
        hg = graphs[g]
        st, moves = R.refine(hg, K, metric, sched, O.build_state(hg, K, p0), passes)
        assert np.array_equal(part, st.part), (g, metric, K)
        ref_m = O.metrics(hg, K, st.part)
        assert info["metrics"] == {k: ref_m[k] for k in O.METRIC_NAMES}, (g, metric, K)
        if moves is not None:
            assert info["moves"] == moves, (g, metric, K)


def test_serve_budget_bounds_the_bytes_in_flight():
    graphs, reqs = _workload()
    out, _ = _serve(graphs, reqs, 1 << 36, 6)
    biggest = max(info["footprint"] for info, _ in out)
    # a budget that fits the largest request once: admission keeps the bytes
    # in flight within it, and the results do not change
    out2, stats = _serve(graphs, reqs, biggest, 6)
    assert stats["peak_bytes"] <= biggest
    for (i1, p1), (i2, p2) in zip(out, out2):
        assert np.array_equal(p1, p2) and i1["metrics"] == i2["metrics"]


def test_replicas_run_one_hypergraphs_requests_side_by_side():
    graphs, reqs = _workload()
    # every request on hypergraph 0, four times over and with more passes, so
    # that requests are still in flight when the next ones arrive
    one = [(0, metric, K, part, sched, 12) for g, metric, K, part, sched, passes in reqs if g == 0] * 4
    with Server(0, hbm_budget=1 << 36, workers=6) as srv:
        srv.set_replicas(3)
        srv.add_graph(graphs[0])
        rids = [srv.submit(g, metric, K, part, sched, passes) for g, metric, K, part, sched, passes in one]
        out = [srv.wait(r) for r in rids]
        stats = srv.stats()
    assert stats["peak_inflight"] > 1  # one hypergraph, several requests at once
    for (info, part), (g, metric, K, p0, sched, passes) in zip(out, one):
        want_part, want_m = _alone(graphs[0], metric, K, p0, sched, passes)
        assert np.array_equal(part, want_part), (metric, K)
        assert info["metrics"] == want_m, (metric, K)
