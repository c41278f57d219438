"""Batched refinement (one HIP graph launch per pass over many hypergraphs)
against the same hypergraphs refined alone through fp_refine: every member's
state, sizes and move count must be identical after every call, including
calls that replay an already captured batch graph with host work and
synchronisations in between."""
import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd.engine import SCHED_BAL, SCHED_REF, Batch, Engine, schedule
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _members():
    """(hypergraph, K, metric, schedule, initial state) of mixed size, K and objective."""
    out = []
    specs = [
        ("TV", H.colnet(3000, H.DEG_UNIFORM, 2, 9, seed=1), 16, "ref", 64),
        ("TV", H.workload("c2", scale=0.01), 64, "bal", 2048),
        ("CN", H.colnet(2500, H.DEG_UNIFORM, 2, 9, seed=2), 16, None, 0),
        ("MSV", H.colnet(3000, H.DEG_UNIFORM, 2, 9, seed=3), 12, "ref", 256),
        ("MSRV", H.colnet(2000, H.DEG_UNIFORM, 2, 9, seed=4), 12, "ref", 512),
        ("TM", H.colnet(1500, H.DEG_UNIFORM, 1, 3, seed=3), 64, "ref", 1 << 30),
        ("MSM", H.colnet(3000, H.DEG_UNIFORM, 1, 3, seed=2), 256, "ref", 128),
        ("TV", H.colnet(800, H.DEG_FIXED, 4, seed=5), 256, "ref", 8),
    ]
    for metric, hg, K, kind, R in specs:
        eps = 0.03 if kind == "bal" else 0.05
        C, max_pw, min_pw = H.ref_constants(hg.nov, K, eps)
        if metric == "TM":
            st = O.tm_init(hg, K, C)
        elif metric == "MSM":
            st = O.msm_init(hg, K, C)
        elif metric == "MSV":
            st = O.msv_ldg(hg, K, C)
        elif metric == "MSRV":
            st = O.msrv_ldg(hg, K, C)
        elif metric == "CN":
            st = O.cn_ldg(hg, K, C)
        else:
            st = O.build_state(hg, K, H.initial_partition(hg, K, C, max_pw, min_pw))
        if kind == "bal":
            sched = schedule(SCHED_BAL, R, C, H.balance_cap(hg.nov, K, eps), 0, H.balance_cap(hg.nov, K, eps))
        elif kind == "ref":
            sched = schedule(SCHED_REF, R, C, max_pw, min_pw)
        else:
            sched = None
        out.append((hg, K, metric, sched, st))
    return out


def _open(hg, K, st):
    eng = Engine(0)
    eng.upload(hg)
    eng.set_state(K, st.part, st.ep_ends, st.ep_parts, st.ep_cnt, st.sizes)
    return eng


def _same(a, b):
    x, y = a.get_state(), b.get_state()
    for f in ("part", "ep_ends", "ep_parts", "ep_cnt", "sizes"):
        assert np.array_equal(getattr(x, f), getattr(y, f)), f


def test_batch_equals_solo_with_tuning():
    ms = _members()
    solo = [_open(hg, K, st) for hg, K, _, _, st in ms]
    inbatch = [_open(hg, K, st) for hg, K, _, _, st in ms]
    try:
        with Batch(0) as b:
            for eng, (_, _, metric, sched, _) in zip(inbatch, ms):
                b.add(eng, metric, sched)
            print(f"batch of {len(ms)}: 5 passes, 3 tuning", flush=True)
            b.refine(5, tune_passes=3)
            for i, (eng, (_, _, metric, sched, _)) in enumerate(zip(solo, ms)):
                stats = eng.refine(metric, 5, sched)
                _same(eng, inbatch[i])
                assert b.member(i)["moves"] == stats.moves, metric
    finally:
        for e in solo + inbatch:
            e.close()


def test_batch_graph_replays_match_solo_pass_by_pass():
    ms = _members()
    solo = [_open(hg, K, st) for hg, K, _, _, st in ms]
    inbatch = [_open(hg, K, st) for hg, K, _, _, st in ms]
    try:
        with Batch(0) as b:
            for eng, (_, _, metric, sched, _) in zip(inbatch, ms):
                b.add(eng, metric, sched)
            for call in range(4):
                # the first call captures; later calls replay the same batch graph
                print(f"batch replay call {call}", flush=True)
                b.refine(1 if call < 3 else 2)
                for i, (eng, (_, _, metric, sched, _)) in enumerate(zip(solo, ms)):
                    eng.refine(metric, 1 if call < 3 else 2, sched)
                    _same(eng, inbatch[i])
    finally:
        for e in solo + inbatch:
            e.close()
