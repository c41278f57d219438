"""Refinement at the full size of the 2M-vertex workload (2,000,000 x
2,000,000 column-net hypergraph, diagonal + 12 uniform off-diagonals per row,
26M pins, vertex weights U{1..8}, K = 1024, 15 passes) for TM, MSM, MSV and
MSRV, and on a sparser 2M-vertex instance (diagonal + 1..3 off-diagonals,
6M pins) on which TM, MSM and MSRV make moves.

Bit-exact against the oracle wherever its refinement finishes in seconds:
- on the 26M-pin instance TM, MSM and MSRV accept no move from the
  streaming initial partition: the oracle's first pass shows it (a pass
  without moves leaves the state, and so every later pass, unchanged), and
  the engine's 15 passes must end on exactly that state; MSV moves, and its
  first 2 passes are compared with the oracle's;
- on the sparse instance TM's first pass (thousands of moves) and MSRV's
  first 3 passes are compared with the oracle's.
All 15 passes are then checked through properties that do not depend on size
(MSM's oracle costs O(K^2) per gaining vertex, minutes at K = 1024, so MSM
is checked by these alone):
- part sizes equal the weighted part histogram of the final part vector;
- every net's list holds exactly the (part, pin count) pairs of its pins,
  with the slots past lambda empty;
- the device metrics equal the oracle's from-scratch metrics;
- TM / MSM: the kept part-connection matrix and message counts equal the
  ones recounted from the final state;
- a second run from the same state gives the same result bit for bit.
"""
import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd.engine import SCHED_REF, Engine, schedule
from oracle import oracle as O

import oracle_refine as R

pytestmark = pytest.mark.gpu

K = 1024
PASSES = 15


def _prep(hg):
    C, max_pw, min_pw = H.ref_constants(int(hg.vwgt.sum()), K)
    part = H.initial_partition(hg, K, C, max_pw, min_pw)
    return hg, part, (C, max_pw, min_pw)


@pytest.fixture(scope="module")
def c3():
    return _prep(H.workload("c3"))


@pytest.fixture(scope="module")
def sparse2m():
    return _prep(H.colnet(2_000_000, H.DEG_UNIFORM, 1, 3, seed=21, weights=(1, 8)))


def _round(metric, hg):
    return hg.nov if metric == "TM" else 16384


def _engine_run(hg, part, metric, consts, passes):
    C, max_pw, min_pw = consts
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_partition(K, part)
        stats = eng.refine(metric, passes, schedule(SCHED_REF, _round(metric, hg), C, max_pw, min_pw))
        st = eng.get_state()
        m = eng.metrics()
        conn = eng.connections() if metric in ("TM", "MSM") else None
    return stats, st, m, conn


_list_entries = R.list_entries


def _with_connections(hg, st, msm, K=K):
    return R.with_connections(hg, st, msm, K)


def _oracle_run(hg, part, metric, consts, passes):
    C, _, _ = consts
    st = O.build_state(hg, K, part)
    st.sizes = st.sizes.astype(np.int32)
    if metric == "TM":
        moves = O.tm_refine_par(hg, _with_connections(hg, st, False), C, passes, hg.nov)
    elif metric == "MSM":
        moves = O.msm_refine_par(hg, _with_connections(hg, st, True), passes, 16384)
    else:
        m = O.metrics(hg, K, part)
        if metric == "MSV":
            st.sv = m["sv_msv"].astype(np.int32)
            moves = O.msv_refine_par(hg, st, C, passes, 16384)
        else:
            st.sv, st.rv = m["sv"].astype(np.int32), m["rv"].astype(np.int32)
            moves = O.msrv_refine_par(hg, st, C, passes, 16384)
    return moves, st


def _same_state(got, st):
    assert np.array_equal(got.part, st.part)
    assert np.array_equal(got.ep_ends, st.ep_ends)
    assert np.array_equal(got.ep_parts, st.ep_parts)
    assert np.array_equal(got.ep_cnt, st.ep_cnt)
    assert np.array_equal(got.sizes, st.sizes.astype(np.int64))


def _check_invariants(hg, st):
    part = st.part
    assert part.min() >= 0 and part.max() < K
    assert np.array_equal(st.sizes, np.bincount(part, weights=hg.vwgt, minlength=K).astype(np.int64))
    # slots past lambda are empty; the lists hold the (part, pin count) pairs
    # of the nets' pins: compared with the oracle's lists of the same part
    # vector in an order-free form
    want = O.build_state(hg, K, part)
    assert np.array_equal(st.ep_ends, want.ep_ends)
    used = np.zeros(hg.pins, bool)
    n, _, _ = _list_entries(hg, st)
    lam = st.ep_ends.astype(np.int64)
    used[np.repeat(hg.net_ptr[:-1].astype(np.int64), lam) + (np.arange(int(lam.sum())) -
                                                           np.repeat(np.cumsum(lam) - lam, lam))] = True
    assert np.all(st.ep_parts[~used] == -1) and np.all(st.ep_cnt[~used] == 0)
    gn, gb, gc = _list_entries(hg, st)
    wn, wb, wc = _list_entries(hg, want)
    assert np.array_equal(np.sort(gn * K * 65536 + gb * 65536 + gc), np.sort(wn * K * 65536 + wb * 65536 + wc))
    return want


def _check_full(hg, metric, stats, st, m, conn):
    want = _check_invariants(hg, st)
    ref = O.metrics(hg, K, st.part)
    for k in O.METRIC_NAMES:
        assert m[k] == ref[k], k
    if conn is not None:
        _with_connections(hg, want, False)
        assert np.array_equal(conn[0].reshape(-1), want.pc) and np.array_equal(conn[1], want.msg)


@pytest.mark.parametrize("metric", ["TM", "MSM", "MSRV"])
def test_c3_no_move_objectives_equal_the_oracle(c3, metric):
    hg, part, consts = c3
    print(f"{metric}: nov={hg.nov} pins={hg.pins} K={K} passes={PASSES}", flush=True)
    moves, st = _oracle_run(hg, part, metric, consts, 1)
    assert moves == 0   # a pass without moves: every later pass is the same
    stats, got, m, conn = _engine_run(hg, part, metric, consts, PASSES)
    assert stats.moves == 0
    _same_state(got, st)
    _check_full(hg, metric, stats, got, m, conn)


def test_c3_msv(c3):
    hg, part, consts = c3
    print(f"MSV: nov={hg.nov} pins={hg.pins} K={K}", flush=True)
    moves, st = _oracle_run(hg, part, "MSV", consts, 2)
    stats, got, _, _ = _engine_run(hg, part, "MSV", consts, 2)
    assert moves > 100 and stats.moves == moves
    _same_state(got, st)
    stats, got, m, conn = _engine_run(hg, part, "MSV", consts, PASSES)
    assert stats.moves >= moves
    _check_full(hg, "MSV", stats, got, m, conn)
    stats2, got2, m2, _ = _engine_run(hg, part, "MSV", consts, PASSES)
    assert stats2.moves == stats.moves and m2 == m and np.array_equal(got2.ep_parts, got.ep_parts)


@pytest.mark.parametrize("metric,oracle_passes,min_moves", [("TM", 1, 1000), ("MSRV", 3, 10), ("MSM", 0, 1)])
def test_sparse_2m_moving_objectives(sparse2m, metric, oracle_passes, min_moves):
    hg, part, consts = sparse2m
    print(f"{metric}: nov={hg.nov} pins={hg.pins} K={K}", flush=True)
    if oracle_passes:
        moves, st = _oracle_run(hg, part, metric, consts, oracle_passes)
        stats, got, _, _ = _engine_run(hg, part, metric, consts, oracle_passes)
        assert stats.moves == moves >= min_moves
        _same_state(got, st)
    stats, got, m, conn = _engine_run(hg, part, metric, consts, PASSES)
    assert stats.moves >= min_moves
    _check_full(hg, metric, stats, got, m, conn)
    print(f"{metric}: {stats.moves} moves in {PASSES} passes, {metric} {m[metric]}", flush=True)
    stats2, got2, m2, _ = _engine_run(hg, part, metric, consts, PASSES)
    assert stats2.moves == stats.moves and m2 == m
    assert np.array_equal(got2.part, got.part) and np.array_equal(got2.ep_parts, got.ep_parts)
