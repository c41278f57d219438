"""MSV and MSRV refinement parity: HIP engine (through the C-ABI) against the
oracle's restatement of partition_par (MGV/partition.cpp:706-1040,
MSRV/partition.cpp:769-1110) on the same seeded inputs. Bit-exact: part
vectors, ordered connectivity lists, counts and part sizes; the volumes are
checked through the from-scratch metrics."""
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


def _run(metric, hg, K, passes, R, ordering=None):
    C, max_pw, min_pw = H.ref_constants(hg.nov, K)
    print(f"{metric}: nov={hg.nov} pins={hg.pins} K={K} passes={passes} R={R} "
          f"ordering={'perm' if ordering is not None else 'id'}", flush=True)
    if metric == "MSV":
        st = O.msv_ldg(hg, K, C)
    else:
        st = O.msrv_ldg(hg, K, C, ordering=ordering)
    with Engine(0) as eng:
        eng.upload(hg, ordering=ordering)
        eng.set_state(K, st.part, st.ep_ends, st.ep_parts, st.ep_cnt)
        stats = eng.refine(metric, passes, schedule(SCHED_REF, R, C, max_pw, min_pw))
        got = eng.get_state()
        m = eng.metrics()
    if metric == "MSV":
        moves = O.msv_refine_par(hg, st, C, passes, R)
    else:
        moves = O.msrv_refine_par(hg, st, C, passes, R, ordering=ordering)
    _same_state(got, st)
    assert stats.moves == moves
    ref = O.metrics(hg, K, st.part)
    assert m[metric] == ref[metric]
    return moves


@pytest.mark.parametrize("R", [256, 1024, 3000])
def test_msv_uniform(R):
    hg = H.colnet(3000, H.DEG_UNIFORM, 2, 9, seed=5)
    assert _run("MSV", hg, 12, 6, R) > 0


@pytest.mark.parametrize("R", [256, 4096])
def test_msv_zipf(R):
    hg = H.workload("c2", scale=0.02)
    assert _run("MSV", hg, 64, 4, R) > 0


@pytest.mark.parametrize("R", [256, 3000])
def test_msrv_uniform(R):
    hg = H.colnet(3000, H.DEG_UNIFORM, 2, 9, seed=5)
    _run("MSRV", hg, 12, 6, R)


def test_msrv_permuted_ordering():
    hg = H.colnet(3000, H.DEG_UNIFORM, 2, 9, seed=6)
    perm = np.random.default_rng(3).permutation(hg.nov).astype(np.int32)
    _run("MSRV", hg, 8, 6, 512, ordering=perm)


def test_msrv_zipf():
    hg = H.workload("c2", scale=0.02)
    _run("MSRV", hg, 64, 4, 2048)


@pytest.mark.parametrize("K", [512, 1024, 2048])
def test_msv_many_parts(K):
    # K <= 1024 takes the scan's batched step (with the move-by-move fallback
    # when many parts tie near the max), K = 2048 the move-by-move scan
    hg = H.colnet(20000, H.DEG_ZIPF, 2, 5000, 2.3, seed=2001)
    assert _run("MSV", hg, K, 3, 4096) > 0


@pytest.mark.parametrize("K", [256, 2048])
def test_msrv_many_parts(K):
    hg = H.colnet(20000, H.DEG_ZIPF, 2, 5000, 2.3, seed=2002)
    _run("MSRV", hg, K, 3, 4096)
