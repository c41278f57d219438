"""The hypergraph upload (fp_upload) from every kind of host memory gives the
same resident instance: pageable arrays large enough to go through the
pinned staging buffers in several pieces, arrays already in pinned memory
(copied directly), no vertex view (transposed on the host), and a second
upload into the same context. Each is checked by the state and all six
metrics after the same refinement, bit for bit."""
import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd.engine import SCHED_BAL, Engine, pinned_copy, schedule

pytestmark = pytest.mark.gpu

K = 64


@pytest.fixture(scope="module")
def instance():
    # 500k vertices, 4.7M pins: net_pins and vtx_nets are 19 MB each, two
    # staging pieces of 16 MB; vtx_ptr and net_ptr (2 MB) go directly
    hg = H.workload("c2", seed=3, scale=0.5)
    assert hg.pins * 4 > 16 << 20
    part = np.random.default_rng(3).integers(0, K, hg.nov).astype(np.int32)
    W = hg.total_weight()
    C, max_pw, _ = H.ref_constants(W, K, 0.03)
    cap = H.balance_cap(W, K, 0.03)
    return hg, part, schedule(SCHED_BAL, 65536, C, cap, 0, cap)


def _run(eng, hg, part, sched, **kw):
    eng.upload(hg, **kw)
    eng.set_partition(K, part)
    eng.refine("TV", 2, sched)
    return eng.get_state(), eng.metrics()


def _same(a, b):
    sa, ma = a
    sb, mb = b
    assert ma == mb
    for f in ("part", "ep_ends", "ep_parts", "ep_cnt", "sizes"):
        assert np.array_equal(getattr(sa, f), getattr(sb, f)), f


def test_upload_from_pinned_pageable_and_transposed_agree(instance):
    hg, part, sched = instance
    eng = Engine(0)
    base = _run(eng, hg, part, sched)
    # the same context again: the staging buffers are reused
    _same(base, _run(eng, hg, part, sched))
    # no vertex view: transposed on the host
    _same(base, _run(Engine(0), hg, part, sched, with_vertex_view=False))
    # every array in pinned memory: copied directly
    hp = pinned_copy(hg)
    assert hp.net_pins.base is not None and np.array_equal(hp.net_pins, hg.net_pins)
    _same(base, _run(Engine(0), hp, part, sched))
    assert base[1]["TV"] > 0
