"""Ragged inputs and the edges of the supported envelope, all six objectives
against the oracle (bit-exact states and metrics): isolated vertices,
single-pin nets, one net holding almost every vertex, K = 1 and 2; and K
beyond what an objective's kernels hold fails with FP_ERR_UNSUPPORTED and
leaves the context usable."""
import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd._native import FastPartError
from fastpartitioner_amd.engine import SCHED_REF, Engine, schedule
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP_ERR_UNSUPPORTED = 5


def _ragged(seed=5, nov=3000, isolated=100):
    """One net over every vertex but the last `isolated` ones, a single-pin
    net on every 7th vertex, and random nets of 2..20 pins; every net's
    source is one of its pins (its first); the last vertices are in no net."""
    rng = np.random.default_rng(seed)
    live = nov - isolated
    nets = [np.arange(live)]
    nets += [np.array([v]) for v in range(0, live, 7)]
    for _ in range(2 * live):
        nets.append(rng.choice(live, size=int(rng.integers(2, 21)), replace=False))
    net_ptr = np.zeros(len(nets) + 1, dtype=np.uint32)
    net_ptr[1:] = np.cumsum([len(n) for n in nets])
    net_pins = np.concatenate(nets).astype(np.uint32)
    src = np.array([n[0] for n in nets], dtype=np.int32)
    return H.Hypergraph.from_nets(nov, net_ptr, net_pins, src)


def _same_state(got, st):
    assert np.array_equal(got.part, st.part)
    assert np.array_equal(got.ep_ends, st.ep_ends)
    assert np.array_equal(got.ep_parts, st.ep_parts)
    assert np.array_equal(got.ep_cnt, st.ep_cnt)
    assert np.array_equal(got.sizes, st.sizes.astype(np.int64))


def _initial(metric, hg, K, C, max_pw, min_pw):
    if metric == "TV":
        return O.tv_ldg(hg, K, C, max_pw, min_pw, par_skip=True)[0]
    if metric == "CN":
        return O.cn_ldg(hg, K, C)
    if metric == "TM":
        return O.tm_init(hg, K, C)
    if metric == "MSM":
        return O.msm_init(hg, K, C)
    if metric == "MSV":
        return O.msv_ldg(hg, K, C)
    return O.msrv_ldg(hg, K, C)


def _oracle_refine(metric, hg, st, C, max_pw, min_pw, passes, R):
    if metric == "TV":
        O.tv_refine_par(hg, st, C, max_pw, min_pw, passes, R)
    elif metric == "CN":
        O.cn_refine(hg, st, passes, serial=False)
    elif metric == "TM":
        O.tm_refine_par(hg, st, C, passes, R)
    elif metric == "MSM":
        O.msm_refine_par(hg, st, passes, R)
    elif metric == "MSV":
        O.msv_refine_par(hg, st, C, passes, R)
    else:
        O.msrv_refine_par(hg, st, C, passes, R)


@pytest.mark.parametrize("K", [1, 2, 16])
@pytest.mark.parametrize("metric", ["TV", "CN", "TM", "MSM", "MSV", "MSRV"])
def test_ragged_matches_oracle(metric, K):
    hg = _ragged()
    C, max_pw, min_pw = H.ref_constants(hg.nov, K)
    R = hg.nov if metric == "TM" else 256
    passes = 3
    print(f"{metric}: ragged nov={hg.nov} noe={hg.noe} pins={hg.pins} K={K} R={R}", flush=True)
    st = _initial(metric, hg, K, C, max_pw, min_pw)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_state(K, st.part, st.ep_ends, st.ep_parts, st.ep_cnt, st.sizes)
        sched = None if metric == "CN" else schedule(SCHED_REF, R, C, max_pw, min_pw)
        stats = eng.refine(metric, passes, sched)
        got = eng.get_state()
        m = eng.metrics(vectors=True)
    _oracle_refine(metric, hg, st, C, max_pw, min_pw, passes, R)
    _same_state(got, st)
    ref = O.metrics(hg, K, st.part)
    for k in O.METRIC_NAMES:
        assert m[k] == ref[k], k
    for k in ("lam", "sv", "rv", "sm", "rm", "sv_msv"):
        assert np.array_equal(m[k], ref[k]), k
    assert stats.moves >= 0


@pytest.mark.parametrize("metric", ["TV", "TM", "MSM", "MSV", "MSRV"])
def test_k_beyond_the_kernels_fails_cleanly(metric):
    hg = H.colnet(20000, H.DEG_UNIFORM, 2, 9, seed=7)
    K = 8192
    C, max_pw, min_pw = H.ref_constants(hg.nov, K)
    assert C > 0
    part = np.arange(hg.nov, dtype=np.int32) % K
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_partition(K, part)
        with pytest.raises(FastPartError) as e:
            eng.refine(metric, 1, schedule(SCHED_REF, 512, C, max_pw, min_pw))
        assert e.value.code == FP_ERR_UNSUPPORTED
        # the metrics still evaluate at this K, and the context takes a new state
        m = eng.metrics()
        assert m == {k: v for k, v in O.metrics(hg, K, part).items() if k in O.METRIC_NAMES}
        K2 = 16
        C2, max2, min2 = H.ref_constants(hg.nov, K2)
        eng.set_partition(K2, part % K2)
        eng.refine("TV", 1, schedule(SCHED_REF, 512, C2, max2, min2))


def test_k_beyond_the_packing_is_rejected():
    hg = H.colnet(2000, H.DEG_UNIFORM, 2, 9, seed=7)
    with Engine(0) as eng:
        eng.upload(hg)
        with pytest.raises(FastPartError):
            eng.set_partition(16385, np.arange(hg.nov, dtype=np.int32))


@pytest.mark.parametrize("mode", ["ldg_bands", "ldg", "cn_ldg"])
@pytest.mark.parametrize("K", [1, 2, 16])
@pytest.mark.parametrize("R", [1, 64, 5000])
def test_device_stream_ragged(mode, K, R):
    # the device initial partition on the ragged hypergraph: isolated
    # vertices (fallback turns), single-pin nets, one net over almost every
    # vertex (its list grows to K entries), a round longer than the graph
    hg = _ragged(seed=6, nov=1500, isolated=60)
    W = hg.total_weight()
    C, max_pw, _ = H.ref_constants(W, K)
    cap = H.balance_cap(W, K, 0.05)
    print(f"device stream ragged: mode={mode} K={K} R={R}", flush=True)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.initial_partition({"ldg_bands": 0, "ldg": 1, "cn_ldg": 2}[mode], K, R, C, max_pw, cap)
        got = eng.get_state()
    want, _ = O.round_ldg(hg, K, mode, C, max_pw, R, cap)
    assert np.array_equal(got.part, want.part)
    assert np.array_equal(got.ep_ends, want.ep_ends)
    assert np.array_equal(got.ep_parts, want.ep_parts)
    assert np.array_equal(got.ep_cnt, want.ep_cnt)
    assert np.array_equal(got.sizes, want.sizes)


@pytest.mark.parametrize("view", ["ascending", "file"])
def test_bfs_ragged(view):
    # isolated vertices are never reached and follow in ascending order; the
    # big net is expanded once, by the first frontier vertex that holds it
    from fastpartitioner_amd.engine import BFS_PINS_ASCENDING, BFS_PINS_UPLOADED
    hg = _ragged(seed=7, nov=2000, isolated=80)
    print(f"bfs ragged: view={view}", flush=True)
    with Engine(0) as eng:
        eng.upload(hg)
        got, f1, f2, st = eng.bfs_ordering(3, BFS_PINS_ASCENDING if view == "ascending" else BFS_PINS_UPLOADED)
    want, g1, g2 = O.bfs_ordering(hg, 3, view)
    assert (f1, f2) == (g1, g2)
    assert np.array_equal(got, want)
    assert st.reached == hg.nov - 80
    assert np.array_equal(got[-80:], np.arange(hg.nov - 80, hg.nov))


def test_device_stream_rejects_bad_arguments():
    hg = _ragged(seed=8, nov=500, isolated=10)
    with Engine(0) as eng:
        eng.upload(hg)
        for mode, K, R, C in ((3, 4, 16, 100.0), (0, 0, 16, 100.0), (0, 4, 0, 100.0), (0, 4, 16, 0.0)):
            with pytest.raises(FastPartError):
                eng.initial_partition(mode, K, R, C, 1000, 1000)
        # the context stays usable
        eng.initial_partition(0, 4, 16, 130.0, 1000, 140)
        assert eng.get_state().part.min() >= 0


@pytest.mark.parametrize("metric", ["TV", "MSV", "TM"])
def test_kernel_timing_times_every_evaluation_and_changes_nothing(metric):
    # the evaluation launches timed by their own dispatch (fp_set_kernel_timing):
    # one measurement per round, and the same refinement as without timing
    hg = H.workload("c2", seed=2, scale=0.02)
    K, passes, R = 64, 3, 4096
    C, max_pw, min_pw = H.ref_constants(hg.nov, K)
    part = H.initial_partition_for(metric, hg, K)
    s = schedule(SCHED_REF, hg.nov if metric == "TM" else R, C, max_pw, min_pw)
    out = []
    for timed in (False, True):
        with Engine(0) as eng:
            eng.upload(hg)
            eng.set_partition(K, part)
            eng.set_kernel_timing(timed)
            st = eng.refine(metric, passes, s)
            launches, ms = eng.kernel_timing()
            out.append((st.moves, eng.get_state().part, eng.metrics()))
            if timed:
                assert launches == st.rounds and ms > 0
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1]) and out[0][2] == out[1][2]
