"""TM and MSM refinement parity: HIP engine (through the C-ABI) against the
oracle's restatement of partition_par (TM/partition.cpp:947-1265,
MSM/partition.cpp:1010-1400) on the same seeded inputs. Bit-exact: part
vectors, ordered connectivity lists, counts, part sizes and applied moves;
TM / MSM through the from-scratch metrics."""
import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd.engine import SCHED_REF, Engine, schedule
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _same_state(got, st):
    assert np.array_equal(got.part, st.part)
    assert np.array_equal(got.ep_ends, st.ep_ends)
    assert np.array_equal(got.ep_parts, st.ep_parts)
    assert np.array_equal(got.ep_cnt, st.ep_cnt)
    assert np.array_equal(got.sizes, st.sizes.astype(np.int64))


def _run(metric, hg, K, passes, R):
    C, max_pw, min_pw = H.ref_constants(hg.nov, K)
    print(f"{metric}: nov={hg.nov} pins={hg.pins} K={K} passes={passes} R={R}", flush=True)
    st = O.tm_init(hg, K, C) if metric == "TM" else O.msm_init(hg, K, C)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_state(K, st.part, st.ep_ends, st.ep_parts, st.ep_cnt, st.sizes)
        stats = eng.refine(metric, passes, schedule(SCHED_REF, R, C, max_pw, min_pw))
        got = eng.get_state()
        m = eng.metrics()
        pc, msg = eng.connections()
    if metric == "TM":
        moves = O.tm_refine_par(hg, st, C, passes, R)
    else:
        moves = O.msm_refine_par(hg, st, passes, R)
    _same_state(got, st)
    assert stats.moves == moves
    assert m[metric] == O.metrics(hg, K, st.part)[metric]
    # the connection matrix the oracle kept move by move (its MSM diagonal is -1)
    want = st.pc.reshape(K, K).copy()
    np.fill_diagonal(want, 0)
    assert np.array_equal(pc, want) and np.array_equal(msg, st.msg)
    return moves


@pytest.mark.parametrize("R", [1 << 30, 512])
def test_tm_sparse(R):
    hg = H.colnet(3000, H.DEG_UNIFORM, 0, 3, seed=3)
    assert _run("TM", hg, 256, 4, R) > 50


def test_tm_k64():
    hg = H.colnet(1500, H.DEG_UNIFORM, 1, 3, seed=3)
    assert _run("TM", hg, 64, 4, 1 << 30) > 50


def test_tm_zipf():
    hg = H.workload("c2", scale=0.02)
    _run("TM", hg, 64, 3, 1 << 30)


@pytest.mark.parametrize("R", [128, 1024])
def test_msm_moves(R):
    hg = H.colnet(3000, H.DEG_UNIFORM, 1, 3, seed=2)
    assert _run("MSM", hg, 256, 6, R) > 0


def test_msm_k64():
    hg = H.colnet(2000, H.DEG_UNIFORM, 1, 4, seed=3)
    _run("MSM", hg, 64, 4, 128)


@pytest.mark.parametrize("metric", ["TM", "MSM"])
def test_conn_lists_beyond_a_wavefront(metric):
    # Zipf nets of up to 5000 pins spread over K = 256 parts: vertices with
    # more than 64 receivers (their rows read in several wavefront loads)
    hg = H.workload("c2", scale=0.02)
    _run(metric, hg, 256, 3, 1 << 30 if metric == "TM" else 512)


@pytest.mark.parametrize("metric", ["TM", "MSM"])
def test_conn_k2048(metric):
    # 32 bitmap words per row; receiver lists sized by the instance, far below K
    hg = H.colnet(6000, H.DEG_UNIFORM, 1, 3, seed=4)
    _run(metric, hg, 2048, 3, 1 << 30 if metric == "TM" else 1024)
