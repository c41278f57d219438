"""TV refinement and metric parity: HIP engine (through the C-ABI) against the
oracle (oracle/fp_oracle.c) on the same seeded inputs. Bit-exact: part
vectors, lambda, the ordered connectivity lists, counts, sizes, gains and all
six metrics."""
import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd.engine import SCHED_BAL, SCHED_REF, Engine, schedule
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ldg_state(hg, K, eps=0.05):
    C, max_pw, min_pw = H.ref_constants(hg.nov, K, eps)
    st, _ = O.tv_ldg(hg, K, C, max_pw, min_pw, par_skip=True)
    return st, C, max_pw, min_pw


def _same_state(got, st):
    assert np.array_equal(got.part, st.part)
    assert np.array_equal(got.ep_ends, st.ep_ends)
    assert np.array_equal(got.ep_parts, st.ep_parts)
    assert np.array_equal(got.ep_cnt, st.ep_cnt)


@pytest.mark.parametrize("R", [8, 64, 1000, 10_000])
def test_tv_ref_schedule_c1(R):
    hg = H.workload("c1")
    K = 16
    st, C, max_pw, min_pw = _ldg_state(hg, K)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_state(K, st.part, st.ep_ends, st.ep_parts, st.ep_cnt)
        eng.refine("TV", 10, schedule(SCHED_REF, R, C, max_pw, min_pw))
        got = eng.get_state()
    O.tv_refine_par(hg, st, C, max_pw, min_pw, 10, R)
    _same_state(got, st)
    assert np.array_equal(got.sizes, st.sizes.astype(np.int64))


@pytest.mark.parametrize("R", [64, 4096])
def test_tv_bal_schedule(R):
    hg = H.workload("c2", scale=0.02)
    K = 64
    eps = 0.03
    C, max_pw, min_pw = H.ref_constants(hg.nov, K, eps)
    cap = H.balance_cap(hg.nov, K, eps)
    st0, _ = O.tv_ldg(hg, K, C, max_pw, min_pw, par_skip=True)
    st = O.build_state(hg, K, st0.part)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_partition(K, st0.part)
        s0 = eng.get_state()
        _same_state(s0, st)
        eng.refine("TV", 5, schedule(SCHED_BAL, R, C, cap, 0, cap))
        got = eng.get_state()
    O.tv_refine_bal(hg, st, C, cap, 0, cap, 5, R)
    _same_state(got, st)
    assert np.array_equal(got.sizes, st.sizes)


@pytest.mark.parametrize("K", [8, 37, 64, 100, 1024, 2048, 4096, 5000, 16384])
def test_metrics_match_oracle(K):
    # K not a multiple of 32 pads the message bitmap's rows; above 4096 the
    # per-part counters live in HBM
    hg = H.workload("c2", scale=0.02)
    rng = np.random.default_rng(3)
    part = rng.integers(0, K, hg.nov).astype(np.int32)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_partition(K, part)
        m = eng.metrics(vectors=True)
        again = eng.metrics()
    r = O.metrics(hg, K, part)
    for k in O.METRIC_NAMES:
        assert m[k] == r[k] == again[k], k
    for k in ("lam", "sv", "rv", "sm", "rm", "sv_msv"):
        assert np.array_equal(m[k], r[k]), k


def _window_edge_hypergraph(seed):
    """Nets whose sizes sit on the metric kernel's 64-pin window edges (1, 2,
    63, 64, 65, 127, 128, 129, 1000 pins, and runs of small nets that pack a
    window exactly), some with a repeated pin, sources among the pins."""
    rng = np.random.default_rng(seed)
    nov = 3000
    sizes = [1, 2, 63, 64, 65, 127, 128, 129, 1000] + list(rng.integers(1, 9, 400)) + [32, 32, 16, 16, 16, 16]
    sizes += [64] * 5 + list(rng.integers(1, 70, 300))
    nets = []
    for z in sizes:
        pins = list(rng.choice(nov, size=int(z), replace=False))
        if z > 3 and rng.random() < 0.1:
            pins[-1] = pins[0]  # a repeated pin
        nets.append(pins)
    ptr = np.cumsum([0] + [len(n) for n in nets]).astype(np.uint32)
    pins = np.array(sum(nets, []), dtype=np.uint32)
    src = np.array([n[int(rng.integers(len(n)))] for n in nets], dtype=np.int32)
    return H.Hypergraph.from_nets(nov, ptr, pins, src)


@pytest.mark.parametrize("K,seed", [(5, 1), (37, 2), (300, 3)])
def test_metrics_window_edges(K, seed):
    hg = _window_edge_hypergraph(seed)
    part = np.random.default_rng(seed).integers(0, K, hg.nov).astype(np.int32)
    print(f"metrics window edges: K={K} seed={seed} noe={hg.noe} pins={hg.pins}", flush=True)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_partition(K, part)
        m = eng.metrics(vectors=True)
    r = O.metrics(hg, K, part)
    for k in O.METRIC_NAMES:
        assert m[k] == r[k], k
    for k in ("lam", "sv", "rv", "sm", "rm", "sv_msv"):
        assert np.array_equal(m[k], r[k]), k


def test_gains_match_oracle():
    hg = H.workload("c1", scale=0.5)
    K = 16
    st, C, max_pw, min_pw = _ldg_state(hg, K)
    gains = np.zeros(hg.nov, np.int32)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_state(K, st.part, st.ep_ends, st.ep_parts, st.ep_cnt)
        dec, gain = eng.eval_moves("TV", schedule(SCHED_REF, hg.nov, C, max_pw, min_pw))
    O.tv_refine_par(hg, st.copy(), C, max_pw, min_pw, 1, hg.nov, gains=gains)
    assert np.array_equal(gain, gains)


@pytest.mark.parametrize("name,scale,K", [("c1", 1.0, 16), ("c2", 0.02, 64)])
def test_cn_refinement_matches_oracle(name, scale, K):
    hg = H.workload(name, scale=scale)
    C, _, _ = H.ref_constants(hg.nov, K)
    st = O.cn_ldg(hg, K, C)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_state(K, st.part, st.ep_ends, st.ep_parts, st.ep_cnt)
        stats = eng.refine("CN", 6, schedule(SCHED_REF, 1, C, 0, 0))
        got = eng.get_state()
        m = eng.metrics()
    O.cn_refine(hg, st, 6, serial=False)
    _same_state(got, st)
    assert np.array_equal(got.sizes, st.sizes.astype(np.int64))
    assert m["CN"] == O.metrics(hg, K, st.part)["CN"]
    assert stats.moves > 0


def test_cn_edit_list_growth_reruns_the_overflowed_passes():
    # vertices in about 20 nets each, pushed in bulk from a random 2-way
    # partition: a pass edits more lists than four times the pins, the edit
    # list's first size. The passes are enqueued back to back; the one that
    # overflows and the ones after it leave the state unchanged and run again
    # once the list has grown, which must give the oracle's refinement
    hg = H.colnet(3000, H.DEG_FIXED, 19, seed=9)
    K = 2
    part = np.random.default_rng(9).integers(0, K, hg.nov).astype(np.int32)
    st = O.build_state(hg, K, part)
    st.sizes = st.sizes.astype(np.int32)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_partition(K, part)
        eng.refine("CN", 0)
        b0 = eng.device_bytes()
        stats = eng.refine("CN", 4)
        b1 = eng.device_bytes()
        got = eng.get_state()
    O.cn_refine(hg, st, 4, serial=False)
    assert b1 > b0  # the edit list grew
    assert stats.moves > 0
    _same_state(got, st)
    assert np.array_equal(got.sizes, st.sizes.astype(np.int64))


def _hubs(seed=11, nov=3000, degrees=(62, 63, 64, 125, 126, 127, 188, 189, 190, 200)):
    """Random nets of 2..12 pins, plus hub vertices 0, 1, ... that belong to
    the given numbers of nets: degrees at and around the ends of the
    evaluation's 63-net chunks (63, 126, 189), so that its net groups reach
    the end of a chunk."""
    rng = np.random.default_rng(seed)
    nets = [list(rng.choice(np.arange(len(degrees), nov), size=int(rng.integers(2, 13)), replace=False))
            for _ in range(3 * nov)]
    for h, d in enumerate(degrees):
        for i in rng.choice(len(nets), size=d, replace=False):
            nets[i].append(h)
    net_ptr = np.zeros(len(nets) + 1, dtype=np.uint32)
    net_ptr[1:] = np.cumsum([len(n) for n in nets])
    net_pins = np.concatenate([np.array(n) for n in nets]).astype(np.uint32)
    src = np.array([n[0] for n in nets], dtype=np.int32)
    return H.Hypergraph.from_nets(nov, net_ptr, net_pins, src)


@pytest.mark.parametrize("sched,R", [(SCHED_REF, 64), (SCHED_REF, 1000), (SCHED_BAL, 256)])
def test_tv_hub_vertices(sched, R):
    hg = _hubs()
    assert int(np.diff(hg.vtx_ptr)[:10].min()) == 62 and int(np.diff(hg.vtx_ptr)[:10].max()) == 200
    K = 16
    st, C, max_pw, min_pw = _ldg_state(hg, K)
    cap = H.balance_cap(hg.nov, K, 0.05)
    if sched == SCHED_BAL:
        st = O.build_state(hg, K, st.part)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_state(K, st.part, st.ep_ends, st.ep_parts, st.ep_cnt)
        if sched == SCHED_REF:
            eng.refine("TV", 6, schedule(SCHED_REF, R, C, max_pw, min_pw))
        else:
            eng.refine("TV", 6, schedule(SCHED_BAL, R, C, cap, 0, cap))
        got = eng.get_state()
    if sched == SCHED_REF:
        O.tv_refine_par(hg, st, C, max_pw, min_pw, 6, R)
    else:
        O.tv_refine_bal(hg, st, C, cap, 0, cap, 6, R)
    _same_state(got, st)
    assert np.array_equal(got.sizes, st.sizes.astype(np.int64))


def _long_lists(seed=5, nov=6000, K=256):
    """Connectivity lists past the first 64-entry window: 40 nets of 400
    random pins (lambda ~ 200 of K = 256 under a random partition, so each
    needs 3 more windows), with ids 0..39 so that they open the net lists of
    the hub vertices 0..9. Hub h belongs to 1, 3 or 5 of them (odd task
    counts in its first 63-net chunk) and to random small nets up to its
    degree (62..200, around the chunk ends)."""
    rng = np.random.default_rng(seed)
    degrees = (62, 63, 64, 125, 126, 127, 188, 189, 190, 200)
    nbig = 40
    nets = [list(rng.choice(np.arange(len(degrees), nov), size=400, replace=False)) for _ in range(nbig)]
    nets += [list(rng.choice(np.arange(len(degrees), nov), size=int(rng.integers(2, 13)), replace=False))
             for _ in range(3 * nov)]
    for h, d in enumerate(degrees):
        big = rng.choice(nbig, size=(1, 3, 5)[h % 3], replace=False)
        small = rng.choice(np.arange(nbig, len(nets)), size=d - len(big), replace=False)
        for i in list(big) + list(small):
            nets[i].append(h)
    net_ptr = np.zeros(len(nets) + 1, dtype=np.uint32)
    net_ptr[1:] = np.cumsum([len(n) for n in nets])
    net_pins = np.concatenate([np.array(n) for n in nets]).astype(np.uint32)
    src = np.array([n[0] for n in nets], dtype=np.int32)
    hg = H.Hypergraph.from_nets(nov, net_ptr, net_pins, src)
    part = rng.integers(0, K, nov).astype(np.int32)
    return hg, part


@pytest.mark.parametrize("sched,R", [(SCHED_REF, 64), (SCHED_REF, 6000), (SCHED_BAL, 512)])
def test_tv_long_lists_k256(sched, R):
    # lists of more than 127 entries: the evaluation's further windows, its
    # task list over long nets and the repeat of the last task of an odd count
    K = 256
    hg, part = _long_lists()
    st = O.build_state(hg, K, part)
    assert int(st.ep_ends[:40].min()) >= 128
    C, max_pw, min_pw = H.ref_constants(hg.nov, K, 0.05)
    cap = H.balance_cap(hg.nov, K, 0.05)
    print(f"long lists: K={K} R={R} sched={sched} max lambda={int(st.ep_ends.max())}", flush=True)
    s = schedule(SCHED_REF, R, C, max_pw, min_pw) if sched == SCHED_REF else schedule(SCHED_BAL, R, C, cap, 0, cap)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_partition(K, part)
        dec, gain = eng.eval_moves("TV", s)
        stats = eng.refine("TV", 4, s)
        got = eng.get_state()
    if sched == SCHED_REF:
        ref = O.build_state(hg, K, part)
        ref.sizes = ref.sizes.astype(np.int32)
        gains = np.zeros(hg.nov, np.int32)
        O.tv_refine_par(hg, ref.copy(), C, max_pw, min_pw, 1, hg.nov, gains=gains)
        assert np.array_equal(gain, gains)
        st.sizes = st.sizes.astype(np.int32)
        O.tv_refine_par(hg, st, C, max_pw, min_pw, 4, R)
    else:
        O.tv_refine_bal(hg, st, C, cap, 0, cap, 4, R)
    assert stats.moves > 0
    _same_state(got, st)
    assert np.array_equal(got.sizes, st.sizes.astype(np.int64))


def test_tv_bal_headline_shape():
    # the headline workload's shape at a tenth of its size: Zipf(2.1) column
    # nets, K = 256, eps = 0.03, rounds of a quarter of the vertices, 20 passes
    hg = H.workload("c2", seed=1, scale=0.1)
    K, eps = 256, 0.03
    W = hg.total_weight()
    C, max_pw, min_pw = H.ref_constants(W, K, eps)
    cap = H.balance_cap(W, K, eps)
    part = H.initial_partition(hg, K, C, max_pw, min_pw)
    R = (hg.nov + 3) // 4
    print(f"bal headline shape: nov={hg.nov} pins={hg.pins} R={R}", flush=True)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_partition(K, part)
        stats = eng.refine("TV", 20, schedule(SCHED_BAL, R, C, cap, 0, cap))
        got = eng.get_state()
    st = O.build_state(hg, K, part)
    O.tv_refine_bal(hg, st, C, cap, 0, cap, 20, R)
    assert stats.moves > 1000
    _same_state(got, st)
    assert np.array_equal(got.sizes, st.sizes)


@pytest.mark.parametrize("scale,K,R,passes", [(0.2, 8, 16, 40), (0.1, 16, 8, 80)])
def test_tv_ref_sizes_below_zero(scale, K, R, passes):
    # the round schedule lowers only the source part's size
    # (TV/partition.cpp:827-829): with the size bands open, a long run drives
    # sizes below zero, where LDG2 still breaks score ties by the smaller
    # size as a signed int (TV/partition.cpp:108)
    hg = H.workload("c1", scale=scale)
    st, C, _, _ = _ldg_state(hg, K)
    big = 1 << 30
    print(f"negative sizes: nov={hg.nov} K={K} R={R} passes={passes}", flush=True)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_state(K, st.part, st.ep_ends, st.ep_parts, st.ep_cnt)
        eng.refine("TV", passes, schedule(SCHED_REF, R, C, big, -big))
        got = eng.get_state()
    O.tv_refine_par(hg, st, C, big, -big, passes, R)
    assert st.sizes.min() < 0
    _same_state(got, st)
    assert np.array_equal(got.sizes, st.sizes.astype(np.int64))


def test_cn_net_with_more_moves_than_an_lds_sort():
    # net 0 holds every vertex and is cut; 5,000 two-pin nets (2i, 2i+1) with
    # their pins in parts 0 and 1 push one pin each, so net 0 receives 5,000
    # list edits in one pass: more than the apply sorts in LDS (4,096)
    nov = 10000
    nets = [list(range(nov))] + [[2 * i, 2 * i + 1] for i in range(nov // 2)]
    ptr = np.cumsum([0] + [len(n) for n in nets]).astype(np.uint32)
    hg = H.Hypergraph.from_nets(nov, ptr, np.concatenate([np.array(n) for n in nets]).astype(np.uint32),
                                np.array([n[0] for n in nets], dtype=np.int32))
    K = 4
    part = np.array([(v % 2) if v < 6000 else 2 + (v % 2) for v in range(nov)], dtype=np.int32)
    st = O.build_state(hg, K, part)
    print(f"cn big net: nov={nov} nets={hg.noe}", flush=True)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_state(K, st.part, st.ep_ends, st.ep_parts, st.ep_cnt)
        stats = eng.refine("CN", 2, schedule(SCHED_REF, 1, 1.0, 0, 0))
        got = eng.get_state()
    st.sizes = st.sizes.astype(np.int32)
    O.cn_refine(hg, st, 2, serial=False)
    assert stats.moves >= 4097
    _same_state(got, st)
    assert np.array_equal(got.sizes, st.sizes.astype(np.int64))
