"""Refinement at the full size of the 2M-vertex workload (2,000,000 x
2,000,000 column-net hypergraph, diagonal + 12 uniform off-diagonals per row,
26M pins, vertex weights U{1..8}, K = 1024) for TM, MSM, MSV and MSRV, where
the CPU oracle's refinement would take too long. Checked through properties
that do not depend on size:
- part sizes equal the weighted part histogram of the final part vector;
- every net's list holds exactly the (part, pin count) pairs of its pins,
  with the slots past lambda empty;
- the device metrics equal the oracle's from-scratch metrics of the final
  part vector;
- TM / MSM: the kept part-connection matrix and message counts equal the
  ones recounted from the final state;
- a second run from the same state gives the same result bit for bit.
"""
import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd.engine import SCHED_REF, Engine, schedule
from oracle import oracle as O

pytestmark = pytest.mark.gpu

K = 1024
PASSES = 2


@pytest.fixture(scope="module")
def c3():
    hg = H.workload("c3")
    C, max_pw, min_pw = H.ref_constants(int(hg.vwgt.sum()), K)
    part = H.initial_partition(hg, K, C, max_pw, min_pw)
    w = int(hg.net_ptr[1] - hg.net_ptr[0])
    assert np.all(np.diff(hg.net_ptr) == w)  # every net has the same pin count: lists fit rows of w slots
    return hg, part, (C, max_pw, min_pw), w


def _refine(hg, part, metric, consts):
    C, max_pw, min_pw = consts
    R = hg.nov if metric == "TM" else 16384
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_partition(K, part)
        stats = eng.refine(metric, PASSES, schedule(SCHED_REF, R, C, max_pw, min_pw))
        st = eng.get_state()
        m = eng.metrics()
        conn = eng.connections() if metric in ("TM", "MSM") else None
    return stats, st, m, conn


def _rows(st, w):
    """Each net's list as a row of w slots, (part << 8 | count) per slot, the
    empty slots as -256, sorted: equal rows are equal (part, count) sets."""
    return np.sort((st.ep_parts.reshape(-1, w) << 8) | st.ep_cnt.reshape(-1, w), axis=1)


def _check_state(hg, st, w):
    part = st.part
    assert part.min() >= 0 and part.max() < K
    want_sizes = np.bincount(part, weights=hg.vwgt, minlength=K).astype(np.int64)
    assert np.array_equal(st.sizes, want_sizes)
    # slots past lambda are empty; the lists hold the (part, pin count) pairs
    # of the nets' pins, checked against the oracle's lists of the same part
    # vector in an order-free form
    lam = st.ep_ends.astype(np.int64)
    dead = np.arange(w)[None, :] >= lam[:, None]
    assert np.all(st.ep_parts.reshape(-1, w)[dead] == -1) and np.all(st.ep_cnt.reshape(-1, w)[dead] == 0)
    want = O.build_state(hg, K, part)
    assert np.array_equal(st.ep_ends, want.ep_ends)
    assert np.array_equal(_rows(st, w), _rows(want, w))
    return want


def _connections(hg, want, w):
    parts = want.ep_parts.reshape(-1, w).astype(np.int64)
    a = want.part[hg.src].astype(np.int64)
    off = (parts >= 0) & (parts != a[:, None])
    pc = np.bincount((a[:, None] * K + parts)[off], minlength=K * K).reshape(K, K).astype(np.int32)
    return pc, (pc > 0).sum(axis=1).astype(np.int32)


@pytest.mark.parametrize("metric", ["TM", "MSM", "MSV", "MSRV"])
def test_full_size_invariants(c3, metric):
    hg, part0, consts, w = c3
    print(f"{metric}: nov={hg.nov} pins={hg.pins} K={K} passes={PASSES}", flush=True)
    stats, st, m, conn = _refine(hg, part0, metric, consts)
    want = _check_state(hg, st, w)
    ref = O.metrics(hg, K, st.part)
    for k in O.METRIC_NAMES:
        assert m[k] == ref[k], k
    if conn is not None:
        pc, msg = _connections(hg, want, w)
        assert np.array_equal(conn[0], pc) and np.array_equal(conn[1], msg)
    print(f"{metric}: moves {stats.moves}, {metric} {ref[metric]}", flush=True)
    stats2, st2, m2, _ = _refine(hg, part0, metric, consts)
    assert stats2.moves == stats.moves and m2 == m
    assert np.array_equal(st2.part, st.part) and np.array_equal(st2.ep_parts, st.ep_parts)
