"""fp_eval_noalloc, the evaluation that allocates nothing (its scratch is a
static device array), against fp_eval on the same states: two calls in a
row, the second on a smaller hypergraph, plus one at K = 1024 (messages
through the global bitmap) and the K limit. Outputs live in device memory the
test allocates through the HIP runtime the engine itself links."""
import ctypes

import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd.engine import Engine

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return _hip


class DeviceBuf:
    def __init__(self, nbytes):
        self.n = nbytes
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), nbytes) == 0

    def fill(self, byte):
        assert hip().hipMemset(self.p, byte, self.n) == 0

    def get(self, dtype, count):
        # the engine's streams do not synchronise with the null stream: wait
        # for the whole device first
        assert hip().hipDeviceSynchronize() == 0
        out = np.empty(count, dtype=dtype)
        assert hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes, 2) == 0  # D2H
        return out

    def free(self):
        hip().hipFree(self.p)


def _expected(eng):
    m = eng.metrics(vectors=True)
    head = [m[k] for k in ("TV", "CN", "TM", "MSM", "MSV", "MSRV")]
    return np.concatenate([np.array(head, np.int64)] + [m[k] for k in ("sv", "rv", "sm", "rm", "sv_msv")]), m["lam"]


def _noalloc(eng, K, noe):
    out = DeviceBuf(8 * (6 + 5 * K))
    lam = DeviceBuf(4 * max(noe, 1))
    out.fill(0x5A)   # the call must overwrite every word
    lam.fill(0x5A)
    before = eng.device_bytes()
    eng._check(eng.lib.fp_eval_noalloc(eng.ctx, out.p, lam.p), "fp_eval_noalloc")
    got = out.get(np.int64, 6 + 5 * K)
    got_lam = lam.get(np.int32, noe)
    assert eng.device_bytes() == before  # nothing allocated in the context
    out.free()
    lam.free()
    return got, got_lam


@pytest.mark.parametrize("first,second", [((0.05, 512), (0.01, 64)), ((0.05, 1024), (0.02, 37))])
def test_noalloc_equals_fp_eval_twice_in_a_row(first, second):
    for scale, K in (first, second):
        hg = H.workload("c2", seed=3, scale=scale)
        part = np.random.default_rng(K).integers(0, K, hg.nov).astype(np.int32)
        print(f"noalloc: nov={hg.nov} noe={hg.noe} K={K}", flush=True)
        with Engine(0) as eng:
            eng.upload(hg)
            eng.set_partition(K, part)
            want, want_lam = _expected(eng)
            got, got_lam = _noalloc(eng, K, hg.noe)
            assert np.array_equal(got, want)
            assert np.array_equal(got_lam, want_lam)


def test_noalloc_rejects_k_above_its_maximum():
    hg = H.colnet(3000, H.DEG_UNIFORM, 2, 9, seed=1)
    K = 2048
    with Engine(0) as eng:
        eng.upload(hg)
        eng.set_partition(K, np.arange(hg.nov, dtype=np.int32) % K)
        out = DeviceBuf(8 * (6 + 5 * K))
        rc = eng.lib.fp_eval_noalloc(eng.ctx, out.p, None)
        out.free()
        assert rc == 5  # FP_ERR_UNSUPPORTED
        eng.metrics()   # the context is still usable
