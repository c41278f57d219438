"""The full-size tests' checkers (tests/test_gpu_scale.py), on the CPU: the
order-free list check and the connection recount must fail on planted
mistakes and pass on the oracle's own states."""
import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from oracle import oracle as O

import test_gpu_scale as S


@pytest.fixture(scope="module")
def small():
    hg = H.colnet(20000, H.DEG_UNIFORM, 1, 3, seed=21, weights=(1, 8))
    part = np.random.default_rng(1).integers(0, S.K, hg.nov).astype(np.int32)
    return hg, part


def test_invariants_hold_on_a_canonical_state(small):
    hg, part = small
    S._check_invariants(hg, O.build_state(hg, S.K, part))


@pytest.mark.parametrize("plant", ["count", "part", "dead_slot", "size"])
def test_invariant_checker_catches_planted_mistakes(small, plant):
    hg, part = small
    st = O.build_state(hg, S.K, part)
    n = int(np.argmax(st.ep_ends >= 2))
    s = int(hg.net_ptr[n])
    if plant == "count":
        st.ep_cnt[s] += 1
    elif plant == "part":
        st.ep_parts[s] = (st.ep_parts[s] + 1) % S.K
    elif plant == "dead_slot":
        st.ep_parts[s + int(st.ep_ends[n]) - 1] = -1
        st.ep_cnt[s + int(st.ep_ends[n]) - 1] = 0
        st.ep_ends[n] -= 1
    else:
        st.sizes[0] += 1
    with pytest.raises(AssertionError):
        S._check_invariants(hg, st)


def test_connection_recount_matches_the_oracle_greedy():
    # the recount of the part-connection matrix equals the matrix the oracle's
    # TM greedy keeps incrementally (TM/partition.cpp:190-235)
    hg = H.colnet(3000, H.DEG_UNIFORM, 1, 3, seed=3)
    K = 64
    C, _, _ = H.ref_constants(hg.nov, K)
    st = O.tm_init(hg, K, C)
    pc, msg = st.pc.copy(), st.msg.copy()
    S._with_connections(hg, st, False, K)
    assert np.array_equal(st.pc, pc) and np.array_equal(st.msg, msg)
