"""Batched refinement (one HIP graph launch per pass over many hypergraphs)
against the oracle's refinement of every member and against the same
hypergraphs refined alone through fp_refine: every member's state, sizes and
move count must be identical after every call, including calls that replay an
already captured batch graph with host work and synchronisations in between,
and calls after a member was re-uploaded or given another K."""
import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd.engine import SCHED_BAL, SCHED_REF, Batch, Engine, schedule
from oracle import oracle as O

import oracle_refine as R

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


def _oracle_refine(hg, K, metric, sched, st, passes):
    return R.refine(hg, K, metric, sched, st, passes)


def _same_as_oracle(eng, st):
    R.assert_same_state(eng.get_state(), st)


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
            for i, (eng, (hg, K, metric, sched, st)) in enumerate(zip(solo, ms)):
                stats = eng.refine(metric, 5, sched)
                _same(eng, inbatch[i])
                assert b.member(i)["moves"] == stats.moves, metric
                want, moves = _oracle_refine(hg, K, metric, sched, st, 5)
                _same_as_oracle(inbatch[i], want)
                if moves is not None:
                    assert b.member(i)["moves"] == moves, metric
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


def _c4_members(seed=4):
    """The batched workload's shape at test size: every objective at every K of
    {16, 64, 256}, column nets with 4..32 uniform off-diagonals plus the
    diagonal, 1,000..6,000 vertices, the batch bench's schedules (TV balanced
    rounds of 16384, TM a whole pass per round, the others reference rounds of
    16384), each objective's own initial partition (the oracle's, which the
    reference pins)."""
    rng = np.random.default_rng(seed)
    out = []
    for metric in ("TV", "CN", "TM", "MSM", "MSV", "MSRV"):
        for K in (16, 64, 256):
            n = int(np.floor(np.exp(rng.uniform(np.log(1e3), np.log(6e3)))))
            hg = H.colnet(n, H.DEG_UNIFORM, 4, 32, seed=int(rng.integers(1 << 30)))
            C, max_pw, min_pw = H.ref_constants(hg.nov, K, 0.05)
            cap = H.balance_cap(hg.nov, K, 0.05)
            if metric == "TV":
                st = O.build_state(hg, K, H.initial_partition(hg, K, C, max_pw, min_pw))
                sched = schedule(SCHED_BAL, 16384, C, cap, 0, cap)
            else:
                st = {"CN": O.cn_ldg, "TM": O.tm_init, "MSM": O.msm_init, "MSV": O.msv_ldg,
                      "MSRV": O.msrv_ldg}[metric](hg, K, C)
                sched = None if metric == "CN" else schedule(SCHED_REF, hg.nov if metric == "TM" else 16384, C,
                                                             max_pw, min_pw)
            out.append((hg, K, metric, sched, st))
    return out


def test_c4_shaped_batch_equals_oracle():
    ms = _c4_members()
    engs = [_open(hg, K, st) for hg, K, _, _, st in ms]
    try:
        with Batch(0) as b:
            b.set_lanes(6)
            for eng, (_, _, metric, sched, _) in zip(engs, ms):
                b.add(eng, metric, sched)
            print(f"c4-shaped batch of {len(ms)}: 2 tuning passes, then 8 captured", flush=True)
            b.refine(2, tune_passes=2)
            b.refine(8)
        for i, (eng, (hg, K, metric, sched, st)) in enumerate(zip(engs, ms)):
            want, _ = _oracle_refine(hg, K, metric, sched, st, 10)
            _same_as_oracle(eng, want)
    finally:
        for e in engs:
            e.close()


def test_batch_after_reupload_and_new_k():
    # a captured batch graph records device pointers, K and the list layout
    # of every member: re-uploading a member, or giving one another K, must
    # not replay a stale graph
    ms = _members()
    engs = [_open(hg, K, st) for hg, K, _, _, st in ms]
    try:
        with Batch(0) as b:
            for eng, (_, _, metric, sched, _) in zip(engs, ms):
                b.add(eng, metric, sched)
            b.refine(2)   # captures
            print("re-upload member 0, K 64 -> 16 on member 1, batch again", flush=True)
            hg0, K0, m0, s0, st0 = ms[0]
            engs[0].upload(hg0)
            engs[0].set_state(K0, st0.part, st0.ep_ends, st0.ep_parts, st0.ep_cnt, st0.sizes)
            hg1, _, m1, _, _ = ms[1]
            K1 = 16
            C1, max1, min1 = H.ref_constants(hg1.nov, K1, 0.03)
            cap1 = H.balance_cap(hg1.nov, K1, 0.03)
            part1 = H.initial_partition(hg1, K1, C1, max1, min1)
            engs[1].set_partition(K1, part1)
            # the batch keeps the member's schedule; its capacity constants
            # are K = 64's, which the oracle below uses as well
            sched1 = ms[1][3]
            b.refine(3)
        want0, _ = _oracle_refine(hg0, K0, m0, s0, st0, 3)
        _same_as_oracle(engs[0], want0)
        st1 = O.build_state(hg1, K1, part1)
        want1, _ = _oracle_refine(hg1, K1, m1, sched1, st1, 3)
        _same_as_oracle(engs[1], want1)
        assert cap1 > 0
    finally:
        for e in engs:
            e.close()
