"""The BFS ordering on the device (fp_bfs_ordering) against the oracle's
one-thread restatement (or_bfs_ordering), which tests/test_oracle.py pins to
the reference's own BFS (shared/utils.cpp, oracle/_ref/ref_bfs) and its
golden fixture. Bit-exact orderings and farthest vertices."""
import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd.engine import BFS_PINS_ASCENDING, BFS_PINS_UPLOADED, Engine
from oracle import oracle as O

pytestmark = pytest.mark.gpu

VIEWS = {"ascending": BFS_PINS_ASCENDING, "file": BFS_PINS_UPLOADED}


def _shuffled_pins(hg, seed):
    rng = np.random.default_rng(seed)
    pins = np.array(hg.net_pins, dtype=np.uint32)
    for n in range(hg.noe):
        s, e = int(hg.net_ptr[n]), int(hg.net_ptr[n + 1])
        rng.shuffle(pins[s:e])
    return H.Hypergraph.from_nets(hg.nov, hg.net_ptr, pins, hg.src)


CASES = {
    "uniform": lambda: H.colnet(6000, H.DEG_UNIFORM, 1, 4, seed=21),
    "shuffled": lambda: _shuffled_pins(H.colnet(6000, H.DEG_UNIFORM, 1, 5, seed=22), 3),
    "zipf": lambda: H.colnet(50000, H.DEG_ZIPF, 2, 800, alpha=2.1, seed=23),
    "sparse": lambda: H.colnet(4000, H.DEG_FIXED, 1, seed=24),
    "single": lambda: H.colnet(1, H.DEG_FIXED, 0, seed=1),
}


def _device(hg, start, view):
    print(f"bfs: nov={hg.nov} pins={hg.pins} start={start} view={view}", flush=True)
    with Engine(0) as eng:
        eng.upload(hg)
        return eng.bfs_ordering(start, VIEWS[view])


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("view", list(VIEWS))
def test_bfs_equals_oracle(case, view):
    hg = CASES[case]()
    start = 1804289383 % hg.nov
    got, f1, f2, st = _device(hg, start, view)
    want, g1, g2 = O.bfs_ordering(hg, start, view)
    assert (f1, f2) == (g1, g2)
    assert np.array_equal(got, want)
    assert 1 <= st.reached <= hg.nov


def test_bfs_golden_instance():
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bfs.npz"))
    hg = H.Hypergraph.from_nets(int(z["nov"]), z["net_ptr"], z["net_pins"], z["src"])
    for view, key in (("ascending", "asc"), ("file", "file")):
        got, f1, f2, _ = _device(hg, int(z["start"]), view)
        ref = z[f"{key}_ordering"]
        n = int(np.count_nonzero(ref >= 0))
        assert (f1, f2) == tuple(int(x) for x in z[f"{key}_farthest"])
        assert np.array_equal(got[:n], ref[:n])


def test_bfs_headline_instance():
    # the 1M-vertex Zipf instance (nets up to 5000 pins)
    hg = H.workload("c2", seed=1)
    start = 1804289383 % hg.nov
    got, f1, f2, st = _device(hg, start, "ascending")
    want, g1, g2 = O.bfs_ordering(hg, start, "ascending")
    assert (f1, f2) == (g1, g2)
    assert np.array_equal(got, want)
    print(f"headline BFS: reached {st.reached}, {st.levels} levels, {st.device_ms:.2f} ms for the three BFS", flush=True)


def test_bfs_ordering_feeds_the_refinement():
    # the ordering goes back to fp_upload; refinement then visits in BFS order
    hg = H.colnet(20000, H.DEG_UNIFORM, 2, 9, seed=31)
    K = 16
    with Engine(0) as eng:
        eng.upload(hg)
        order, _, _, _ = eng.bfs_ordering(5, BFS_PINS_ASCENDING)
    C, max_pw, min_pw = H.ref_constants(hg.nov, K)
    st, _ = O.tv_ldg(hg, K, C, max_pw, min_pw, par_skip=True, ordering=order)
    from fastpartitioner_amd.engine import SCHED_REF, schedule
    with Engine(0) as eng:
        eng.upload(hg, ordering=order)
        eng.set_state(K, st.part, st.ep_ends, st.ep_parts, st.ep_cnt)
        eng.refine("TV", 3, schedule(SCHED_REF, 64, C, max_pw, min_pw))
        got = eng.get_state()
    O.tv_refine_par(hg, st, C, max_pw, min_pw, 3, 64, ordering=order)
    assert np.array_equal(got.part, st.part)
    assert np.array_equal(got.ep_parts, st.ep_parts)
