"""Python side of the HIP engine (include/fastpart.h): one Engine = one
fp_ctx = one hypergraph and its connectivity state resident on one GPU.

The methods mirror the reference's flow around partition_gpu_wrapper
(TV/partition.cpp:984-1184): upload both CSR views once, hand over the state
left by the host-side initial partition, run REFI_ITER refinement passes on
the device, read the refined state and metrics back.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import _native
from .hypergraph import Hypergraph

_U32P = ctypes.POINTER(ctypes.c_uint32)
_I32P = ctypes.POINTER(ctypes.c_int32)
_I64P = ctypes.POINTER(ctypes.c_int64)

METRICS = {"CN": 0, "TV": 1, "TM": 2, "MSM": 3, "MSV": 4, "MSRV": 5}
SCHED_REF, SCHED_BAL = 0, 1


class Schedule(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("round", ctypes.c_int), ("ldg_C", ctypes.c_double),
                ("max_pw", ctypes.c_longlong), ("min_pw", ctypes.c_longlong), ("cap", ctypes.c_longlong)]


class RefineStats(ctypes.Structure):
    _fields_ = [("moves", ctypes.c_longlong), ("rounds", ctypes.c_longlong),
                ("pin_visits", ctypes.c_longlong), ("device_ms", ctypes.c_double)]


class InitStats(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_longlong), ("subrounds", ctypes.c_longlong), ("device_ms", ctypes.c_double)]


class BfsStats(ctypes.Structure):
    _fields_ = [("reached", ctypes.c_longlong), ("levels", ctypes.c_longlong), ("device_ms", ctypes.c_double)]


BFS_PINS_ASCENDING, BFS_PINS_UPLOADED = 0, 1


class Metrics(ctypes.Structure):
    _fields_ = [(k, ctypes.c_longlong) for k in ("tv", "cn", "tm", "msm", "msv", "msrv")]


class RequestInfo(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("graph", ctypes.c_int), ("metric", ctypes.c_int), ("K", ctypes.c_int),
                ("passes", ctypes.c_int), ("status", ctypes.c_int), ("moves", ctypes.c_longlong),
                ("metrics", Metrics), ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("end_ms", ctypes.c_double), ("footprint", ctypes.c_longlong)]


def declare(lib) -> None:
    c = ctypes
    P = c.c_void_p
    lib.fp_ctx_last_error.restype = c.c_char_p
    lib.fp_ctx_last_error.argtypes = [P]
    lib.fp_ctx_device_bytes.restype = c.c_longlong
    lib.fp_ctx_device_bytes.argtypes = [P]
    lib.fp_ctx_create.argtypes = [c.POINTER(P), c.c_int]
    lib.fp_ctx_destroy.argtypes = [P]
    lib.fp_ctx_destroy.restype = None
    lib.fp_device_count.argtypes = [c.POINTER(c.c_int)]
    lib.fp_upload.argtypes = [P, c.c_int, c.c_int, _U32P, _U32P, _U32P, _U32P, _I32P, _I32P, _I32P]
    lib.fp_host_alloc_pinned.restype = P
    lib.fp_host_alloc_pinned.argtypes = [c.c_size_t]
    lib.fp_host_free_pinned.restype = None
    lib.fp_host_free_pinned.argtypes = [P]
    lib.fp_set_partition.argtypes = [P, c.c_int, _I32P]
    lib.fp_bfs_ordering.argtypes = [P, c.c_int, c.c_int, _I32P, c.POINTER(c.c_int), c.POINTER(c.c_int),
                                     c.POINTER(BfsStats)]
    lib.fp_initial_partition_device.argtypes = [P, c.c_int, c.c_int, c.c_int, c.c_double, c.c_longlong,
                                                 c.c_longlong, c.POINTER(InitStats)]
    lib.fp_set_state.argtypes = [P, c.c_int, _I32P, _I32P, _I32P, _I32P, _I64P]
    lib.fp_get_state.argtypes = [P, _I32P, _I32P, _I32P, _I32P, _I64P]
    lib.fp_refine.argtypes = [P, c.c_int, c.c_int, c.POINTER(Schedule), c.POINTER(RefineStats)]
    lib.fp_eval_moves.argtypes = [P, c.c_int, c.POINTER(Schedule), _I32P, _I32P]
    lib.fp_eval.argtypes = [P, c.POINTER(Metrics), _I32P, _I64P, _I64P, _I64P, _I64P, _I64P]
    lib.fp_eval_noalloc.argtypes = [P, P, P]
    lib.fp_get_connections.argtypes = [P, _I32P, _I32P]
    lib.fp_status_string.restype = c.c_char_p
    lib.fp_status_string.argtypes = [c.c_int]
    lib.fp_batch_create.argtypes = [c.POINTER(P), c.c_int]
    lib.fp_batch_destroy.argtypes = [P]
    lib.fp_batch_destroy.restype = None
    lib.fp_batch_last_error.argtypes = [P]
    lib.fp_batch_last_error.restype = c.c_char_p
    lib.fp_batch_add.argtypes = [P, P, c.c_int, c.POINTER(Schedule)]
    lib.fp_batch_size.argtypes = [P]
    lib.fp_batch_set_lanes.argtypes = [P, c.c_int]
    lib.fp_batch_refine.argtypes = [P, c.c_int, c.c_int, c.POINTER(c.c_double)]
    lib.fp_batch_member.argtypes = [P, c.c_int, c.POINTER(c.c_longlong), c.POINTER(c.c_int), c.POINTER(c.c_int),
                                    c.POINTER(c.c_double)]
    lib.fp_server_create.argtypes = [c.POINTER(P), c.c_int, c.c_longlong, c.c_int]
    lib.fp_server_destroy.argtypes = [P]
    lib.fp_server_destroy.restype = None
    lib.fp_server_last_error.argtypes = [P]
    lib.fp_server_last_error.restype = c.c_char_p
    lib.fp_server_set_replicas.argtypes = [P, c.c_int]
    lib.fp_server_add_graph.argtypes = [P, c.c_int, c.c_int, _U32P, _U32P, _I32P, _I32P, c.POINTER(c.c_int)]
    lib.fp_server_submit.argtypes = [P, c.c_int, c.c_int, c.c_int, _I32P, c.POINTER(Schedule), c.c_int,
                                     c.POINTER(c.c_longlong)]
    lib.fp_server_wait.argtypes = [P, c.c_longlong, c.POINTER(RequestInfo), _I32P]
    lib.fp_server_stats.argtypes = [P, c.POINTER(c.c_longlong), c.POINTER(c.c_int), c.POINTER(c.c_longlong)]
    lib.fp_set_kernel_timing.argtypes = [P, c.c_int]
    lib.fp_last_kernel_timing.argtypes = [P, c.POINTER(c.c_longlong), c.POINTER(c.c_double)]


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def device_count() -> int:
    n = ctypes.c_int(0)
    _native.engine().fp_device_count(ctypes.byref(n))
    return n.value


def schedule(kind: int, round_size: int, C: float, max_pw: int, min_pw: int, cap: int = 0) -> Schedule:
    return Schedule(kind, round_size, C, max_pw, min_pw, cap)


class _PinnedBuffer:
    """Page-locked host memory (fp_host_alloc_pinned), freed when the last
    array over it goes."""

    def __init__(self, n: int, dtype):
        self.lib = _native.engine()
        dt = np.dtype(dtype)
        self.ptr = self.lib.fp_host_alloc_pinned(max(n * dt.itemsize, 1))
        if not self.ptr:
            raise MemoryError(f"fp_host_alloc_pinned({n * dt.itemsize}) failed")
        self.__array_interface__ = {"shape": (n,), "typestr": dt.str, "data": (self.ptr, False), "version": 3}

    def __del__(self):
        if getattr(self, "ptr", None):
            self.lib.fp_host_free_pinned(self.ptr)
            self.ptr = None


def pinned_empty(n: int, dtype) -> np.ndarray:
    """An uninitialised 1-D array in pinned host memory: fp_upload copies it
    by direct DMA instead of through its staging buffers."""
    return np.asarray(_PinnedBuffer(n, dtype))


def pinned_copy(hg: Hypergraph) -> Hypergraph:
    """The hypergraph with every array in pinned host memory."""
    def pin(a):
        if a is None:
            return None
        out = pinned_empty(a.shape[0], a.dtype)
        out[:] = a
        return out

    return Hypergraph(hg.nov, hg.noe, pin(hg.net_ptr), pin(hg.net_pins), pin(hg.vtx_ptr), pin(hg.vtx_nets),
                      pin(hg.src), pin(hg.vwgt))


@dataclasses.dataclass
class State:
    part: np.ndarray
    ep_ends: np.ndarray
    ep_parts: np.ndarray
    ep_cnt: np.ndarray
    sizes: np.ndarray


class Engine:
    def __init__(self, device: int = 0):
        self.lib = _native.engine()
        self.ctx = ctypes.c_void_p()
        rc = self.lib.fp_ctx_create(ctypes.byref(self.ctx), device)
        _native.check(rc, "fp_ctx_create")
        self.hg: Hypergraph | None = None
        self.K = 0

    def close(self):
        if self.ctx:
            self.lib.fp_ctx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            detail = self.lib.fp_ctx_last_error(self.ctx)
            _native.check(rc, what, detail.decode() if detail else "")

    def device_bytes(self) -> int:
        return int(self.lib.fp_ctx_device_bytes(self.ctx))

    def upload(self, hg: Hypergraph, ordering=None, with_vertex_view: bool = True):
        self.hg = hg
        ordp = None if ordering is None else np.ascontiguousarray(ordering, dtype=np.int32)
        self._check(self.lib.fp_upload(self.ctx, hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.net_pins, _U32P),
                                       _p(hg.vtx_ptr, _U32P) if with_vertex_view else None,
                                       _p(hg.vtx_nets, _U32P) if with_vertex_view else None,
                                       _p(hg.src, _I32P), _p(hg.vwgt, _I32P), _p(ordp, _I32P)), "fp_upload")

    def set_partition(self, K: int, part):
        self.K = K
        part = np.ascontiguousarray(part, dtype=np.int32)
        self._check(self.lib.fp_set_partition(self.ctx, K, _p(part, _I32P)), "fp_set_partition")

    def bfs_ordering(self, start: int, view: int = BFS_PINS_ASCENDING):
        """The reference's BFS ordering on the device (fp_bfs_ordering):
        (ordering, farthest_first, farthest_all, BfsStats)."""
        order = np.zeros(self.hg.nov, dtype=np.int32)
        f1, f2, st = ctypes.c_int(), ctypes.c_int(), BfsStats()
        self._check(self.lib.fp_bfs_ordering(self.ctx, start, view, _p(order, _I32P), ctypes.byref(f1),
                                             ctypes.byref(f2), ctypes.byref(st)), "fp_bfs_ordering")
        return order, f1.value, f2.value, st

    def initial_partition(self, mode: int, K: int, round: int, C: float, max_pw: int, cap: int) -> InitStats:
        """The streaming initial partition on the device (fp_initial_partition_device):
        mode 0 LDG2 with the max_pw band, 1 band-free LDG, 2 CN's LDG over
        uncut nets; rounds of `round` ordering positions (1 = the reference's
        sequential stream). Leaves the state resident."""
        self.K = K
        st = InitStats()
        self._check(self.lib.fp_initial_partition_device(self.ctx, mode, K, round, C, max_pw, cap, ctypes.byref(st)),
                    "fp_initial_partition_device")
        return st

    def set_state(self, K: int, part, ep_ends, ep_parts, ep_cnt, sizes=None):
        self.K = K
        arrs = [np.ascontiguousarray(x, dtype=np.int32) for x in (part, ep_ends, ep_parts, ep_cnt)]
        sz = None if sizes is None else np.ascontiguousarray(sizes, dtype=np.int64)
        self._check(self.lib.fp_set_state(self.ctx, K, *(_p(x, _I32P) for x in arrs), _p(sz, _I64P)),
                    "fp_set_state")

    def get_state(self) -> State:
        hg = self.hg
        st = State(np.zeros(hg.nov, np.int32), np.zeros(hg.noe, np.int32), np.zeros(hg.pins, np.int32),
                   np.zeros(hg.pins, np.int32), np.zeros(self.K, np.int64))
        self._check(self.lib.fp_get_state(self.ctx, _p(st.part, _I32P), _p(st.ep_ends, _I32P),
                                          _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P), _p(st.sizes, _I64P)),
                    "fp_get_state")
        return st

    def refine(self, metric: str, passes: int, sched: Schedule | None = None) -> RefineStats:
        """`passes` passes of `metric`; sched is required except for CN."""
        stats = RefineStats()
        self._check(self.lib.fp_refine(self.ctx, METRICS[metric], passes,
                                       ctypes.byref(sched) if sched is not None else None,
                                       ctypes.byref(stats)), "fp_refine")
        return stats

    def eval_moves(self, metric: str, sched: Schedule):
        dec = np.zeros(self.hg.nov, np.int32)
        gain = np.zeros(self.hg.nov, np.int32)
        self._check(self.lib.fp_eval_moves(self.ctx, METRICS[metric], ctypes.byref(sched), _p(dec, _I32P),
                                           _p(gain, _I32P)), "fp_eval_moves")
        return dec, gain

    def metrics(self, vectors: bool = False):
        m = Metrics()
        K = self.K
        lam = np.zeros(self.hg.noe, np.int32) if vectors else None
        vec = {k: np.zeros(K, np.int64) for k in ("sv", "rv", "sm", "rm", "sv_msv")} if vectors else {}
        self._check(self.lib.fp_eval(self.ctx, ctypes.byref(m), _p(lam, _I32P), *(
            _p(vec.get(k), _I64P) for k in ("sv", "rv", "sm", "rm", "sv_msv"))), "fp_eval")
        res = {"TV": m.tv, "CN": m.cn, "TM": m.tm, "MSM": m.msm, "MSV": m.msv, "MSRV": m.msrv}
        if vectors:
            res["lam"] = lam
            res.update(vec)
        return res

    def connections(self):
        """(pc[K, K], msg[K]) of the resident state: nets with source in part a and pins in part b."""
        K = self.K
        pc = np.zeros(K * K, np.int32)
        msg = np.zeros(K, np.int32)
        self._check(self.lib.fp_get_connections(self.ctx, _p(pc, _I32P), _p(msg, _I32P)), "fp_get_connections")
        return pc.reshape(K, K), msg

    def set_kernel_timing(self, on: bool):
        self._check(self.lib.fp_set_kernel_timing(self.ctx, int(on)), "fp_set_kernel_timing")

    def kernel_timing(self):
        n, ms = ctypes.c_longlong(), ctypes.c_double()
        self._check(self.lib.fp_last_kernel_timing(self.ctx, ctypes.byref(n), ctypes.byref(ms)),
                    "fp_last_kernel_timing")
        return n.value, ms.value


class Batch:
    """Many resident hypergraphs (Engine contexts) refined together: one HIP
    graph launch per pass, launch shapes picked by the first tune passes."""

    def __init__(self, device: int = 0):
        self.lib = _native.engine()
        declare(self.lib)
        self.b = ctypes.c_void_p()
        _native.check(self.lib.fp_batch_create(ctypes.byref(self.b), device), "fp_batch_create")
        self.members = []

    def _check(self, rc, what):
        if rc != 0:
            raise _native.FastPartError(rc, f"{what}: {self.lib.fp_batch_last_error(self.b).decode()}")

    def add(self, eng: "Engine", metric: str, sched: Schedule | None = None):
        self._check(self.lib.fp_batch_add(self.b, eng.ctx, METRICS[metric],
                                          ctypes.byref(sched) if sched is not None else None), "fp_batch_add")
        self.members.append(eng)

    def set_lanes(self, lanes: int):
        self._check(self.lib.fp_batch_set_lanes(self.b, lanes), "fp_batch_set_lanes")

    def refine(self, passes: int, tune_passes: int = 0) -> float:
        ms = ctypes.c_double()
        self._check(self.lib.fp_batch_refine(self.b, passes, tune_passes, ctypes.byref(ms)), "fp_batch_refine")
        return ms.value

    def member(self, i: int):
        mv, w, g, t = ctypes.c_longlong(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._check(self.lib.fp_batch_member(self.b, i, ctypes.byref(mv), ctypes.byref(w), ctypes.byref(g),
                                             ctypes.byref(t)), "fp_batch_member")
        return {"moves": mv.value, "eval_waves": w.value, "eval_grid": g.value, "tuned_pass_ms": t.value}

    def close(self):
        if self.b:
            self.lib.fp_batch_destroy(self.b)
            self.b = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


METRIC_NAMES = {v: k for k, v in METRICS.items()}


class Server:
    """Serving loop over resident hypergraphs (fp_server_*): submit requests,
    wait for them; workers refine admitted requests concurrently."""

    def __init__(self, device: int = 0, hbm_budget: int = 64 << 30, workers: int = 8):
        self.lib = _native.engine()
        declare(self.lib)
        self.s = ctypes.c_void_p()
        _native.check(self.lib.fp_server_create(ctypes.byref(self.s), device, hbm_budget, workers), "fp_server_create")
        self.graphs = []

    def _check(self, rc, what):
        if rc != 0:
            raise _native.FastPartError(rc, what, self.lib.fp_server_last_error(self.s).decode())

    def set_replicas(self, max_replicas: int) -> None:
        self._check(self.lib.fp_server_set_replicas(self.s, max_replicas), "fp_server_set_replicas")

    def add_graph(self, hg: Hypergraph) -> int:
        gid = ctypes.c_int()
        self._check(self.lib.fp_server_add_graph(self.s, hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.net_pins, _U32P),
                                                 _p(hg.src, _I32P), _p(hg.vwgt, _I32P), ctypes.byref(gid)),
                    "fp_server_add_graph")
        self.graphs.append(hg)
        return gid.value

    def submit(self, graph: int, metric: str, K: int, part, sched: Schedule | None, passes: int) -> int:
        part = np.ascontiguousarray(part, dtype=np.int32)
        rid = ctypes.c_longlong()
        self._check(self.lib.fp_server_submit(self.s, graph, METRICS[metric], K, _p(part, _I32P),
                                              ctypes.byref(sched) if sched is not None else None, passes,
                                              ctypes.byref(rid)), "fp_server_submit")
        return rid.value

    def wait(self, rid: int, with_part: bool = True):
        """(info dict, final part vector or None); raises on a failed request."""
        info = RequestInfo()
        part = None
        # the hypergraph is known only from the info; allocate for the largest
        if with_part:
            part = np.zeros(max(g.nov for g in self.graphs), np.int32)
        rc = self.lib.fp_server_wait(self.s, rid, ctypes.byref(info), _p(part, _I32P))
        self._check(rc, f"request {rid}")
        m = info.metrics
        d = {"id": info.id, "graph": info.graph, "metric": METRIC_NAMES[info.metric], "K": info.K,
             "passes": info.passes, "moves": info.moves, "submit_ms": info.submit_ms, "start_ms": info.start_ms,
             "end_ms": info.end_ms, "footprint": info.footprint,
             "metrics": {"TV": m.tv, "CN": m.cn, "TM": m.tm, "MSM": m.msm, "MSV": m.msv, "MSRV": m.msrv}}
        if with_part:
            part = part[:self.graphs[info.graph].nov].copy()
        return d, part

    def stats(self):
        c, p, b = ctypes.c_longlong(), ctypes.c_int(), ctypes.c_longlong()
        self._check(self.lib.fp_server_stats(self.s, ctypes.byref(c), ctypes.byref(p), ctypes.byref(b)),
                    "fp_server_stats")
        return {"completed": c.value, "peak_inflight": p.value, "peak_bytes": b.value}

    def close(self):
        if self.s:
            self.lib.fp_server_destroy(self.s)
            self.s = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
