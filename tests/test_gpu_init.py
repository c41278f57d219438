"""The streaming initial partition on the device (fp_initial_partition_device)
against the oracle's buffered round stream (or_round_ldg), whose one-vertex
rounds are the reference's sequential streams (tests/test_oracle.py).
Bit-exact: part vectors, the ordered connectivity lists, counts and sizes."""
import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd.engine import SCHED_BAL, Engine, schedule
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MODES = {"ldg_bands": 0, "ldg": 1, "cn_ldg": 2}


def _constants(hg, K, eps=0.05):
    W = hg.total_weight()
    C, max_pw, min_pw = H.ref_constants(W, K, eps)
    return C, max_pw, min_pw, H.balance_cap(W, K, eps)


def _device(hg, K, mode, R, C, max_pw, cap, ordering=None):
    print(f"device stream: nov={hg.nov} pins={hg.pins} K={K} mode={mode} R={R}", flush=True)
    with Engine(0) as eng:
        eng.upload(hg, ordering=ordering)
        stats = eng.initial_partition(MODES[mode], K, R, C, max_pw, cap)
        return eng.get_state(), stats


def _same(got, want):
    for k in ("part", "ep_ends", "ep_parts", "ep_cnt"):
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    assert np.array_equal(got.sizes, want.sizes)


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("R", [1, 5, 97, 1024, 5000])
@pytest.mark.parametrize("K", [7, 64])
def test_device_stream_equals_oracle(mode, R, K):
    hg = H.colnet(3000, H.DEG_UNIFORM, 2, 9, seed=11)
    C, max_pw, _, cap = _constants(hg, K)
    got, stats = _device(hg, K, mode, R, C, max_pw, cap)
    want, sub = O.round_ldg(hg, K, mode, C, max_pw, R, cap)
    _same(got, want)
    assert stats.rounds == (hg.nov + min(R, hg.nov) - 1) // min(R, hg.nov)
    assert stats.subrounds >= sub


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("R", [1, 64, 2048])
def test_device_stream_weighted_and_permuted(mode, R):
    hg = H.colnet(2500, H.DEG_UNIFORM, 1, 12, seed=12, weights=(1, 8))
    K = 24
    C, max_pw, _, cap = _constants(hg, K)
    ordering = np.random.default_rng(5).permutation(hg.nov).astype(np.int32)
    got, _ = _device(hg, K, mode, R, C, max_pw, cap, ordering=ordering)
    want, _ = O.round_ldg(hg, K, mode, C, max_pw, R, cap, ordering=ordering)
    _same(got, want)


@pytest.mark.parametrize("K", [1024, 2048])
def test_device_stream_many_parts(K):
    hg = H.colnet(20000, H.DEG_ZIPF, 2, 300, alpha=2.1, seed=13)
    C, max_pw, _, cap = _constants(hg, K)
    for mode in MODES:
        got, _ = _device(hg, K, mode, 4096, C, max_pw, cap)
        want, _ = O.round_ldg(hg, K, mode, C, max_pw, 4096, cap)
        _same(got, want)


@pytest.mark.parametrize("mode", list(MODES))
def test_device_stream_headline_instance(mode):
    # the 1M-vertex Zipf instance (nets up to 5000 pins), K = 256, rounds of 16384
    hg = H.workload("c2", seed=1)
    K = 256
    C, max_pw, _, cap = _constants(hg, K, 0.03)
    got, stats = _device(hg, K, mode, 16384, C, max_pw, cap)
    want, sub = O.round_ldg(hg, K, mode, C, max_pw, 16384, cap)
    _same(got, want)
    print(f"{mode}: {stats.subrounds} sub-rounds ({sub} needed) in {stats.device_ms:.2f} ms", flush=True)


def test_refinement_continues_from_the_device_stream():
    hg = H.colnet(30000, H.DEG_UNIFORM, 2, 12, seed=14)
    K = 64
    C, max_pw, _, cap = _constants(hg, K)
    sched = schedule(SCHED_BAL, 4096, C, cap, 0, cap)
    with Engine(0) as eng:
        eng.upload(hg)
        eng.initial_partition(MODES["ldg_bands"], K, 4096, C, max_pw, cap)
        mv = eng.refine("TV", 3, sched).moves
        got = eng.get_state()
    st, _ = O.round_ldg(hg, K, "ldg_bands", C, max_pw, 4096, cap)
    O.tv_refine_bal(hg, st, C, cap, 0, cap, 3, 4096)
    assert mv > 0
    _same(got, st)
