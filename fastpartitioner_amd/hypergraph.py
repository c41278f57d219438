"""Hypergraphs as the engine sees them: both CSR views plus net sources and
vertex weights, held as numpy arrays on the host.

The net->pins view is what the reference reader produces
(shared/hypergraph_rw.cpp:97-242); the vertex->nets view is its transpose with
ascending nets per vertex (TV/tv.cpp:76-100). Generators for the benchmark
workloads are seeded and implemented natively (libfastpart_host.so), so the
same instance comes out on any machine and any thread count.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import struct

import numpy as np

from . import _native

_U32P = ctypes.POINTER(ctypes.c_uint32)
_I32P = ctypes.POINTER(ctypes.c_int32)


def _ptr(a, t):
    return a.ctypes.data_as(t)


@dataclasses.dataclass
class Hypergraph:
    nov: int
    noe: int
    net_ptr: np.ndarray    # uint32[noe+1]
    net_pins: np.ndarray   # uint32[pins]
    vtx_ptr: np.ndarray    # uint32[nov+1]
    vtx_nets: np.ndarray   # uint32[pins]
    src: np.ndarray        # int32[noe], source vertex of each net
    vwgt: np.ndarray | None = None  # int32[nov] or None (unit weights)

    @property
    def pins(self) -> int:
        return int(self.net_ptr[-1])

    def total_weight(self) -> int:
        return int(self.nov if self.vwgt is None else int(self.vwgt.sum(dtype=np.int64)))

    @staticmethod
    def from_nets(nov: int, net_ptr, net_pins, src=None, vwgt=None) -> "Hypergraph":
        net_ptr = np.ascontiguousarray(net_ptr, dtype=np.uint32)
        net_pins = np.ascontiguousarray(net_pins, dtype=np.uint32)
        noe = len(net_ptr) - 1
        vtx_ptr, vtx_nets = transpose(nov, net_ptr, net_pins)
        if src is None:
            src = np.full(noe, -1, dtype=np.int32)
        return Hypergraph(nov, noe, net_ptr, net_pins, vtx_ptr, vtx_nets,
                          np.ascontiguousarray(src, dtype=np.int32),
                          None if vwgt is None else np.ascontiguousarray(vwgt, dtype=np.int32))

    def write_fphg(self, path: str) -> None:
        """Binary instance file: b'FPHG', i32 version, i32 nov, i32 noe,
        i64 pins, u32 net_ptr[noe+1], u32 net_pins[pins], i32 src[noe]."""
        with open(path, "wb") as f:
            f.write(b"FPHG")
            f.write(struct.pack("<iiiq", 1, self.nov, self.noe, self.pins))
            f.write(self.net_ptr.astype("<u4").tobytes())
            f.write(self.net_pins.astype("<u4").tobytes())
            f.write(self.src.astype("<i4").tobytes())


def transpose(nov: int, net_ptr: np.ndarray, net_pins: np.ndarray):
    lib = _native.host()
    noe = len(net_ptr) - 1
    vtx_ptr = np.zeros(nov + 1, dtype=np.uint32)
    vtx_nets = np.zeros(len(net_pins), dtype=np.uint32)
    rc = lib.fp_csr_transpose(nov, noe, _ptr(net_ptr, _U32P), _ptr(net_pins, _U32P),
                              _ptr(vtx_ptr, _U32P), _ptr(vtx_nets, _U32P))
    _native.check(rc, "fp_csr_transpose")
    return vtx_ptr, vtx_nets


DEG_FIXED, DEG_UNIFORM, DEG_ZIPF = 0, 1, 2


def colnet(n: int, degree_mode: int, deg_lo: int, deg_hi: int = 0, alpha: float = 0.0, seed: int = 1,
           weights: tuple[int, int] | None = None) -> Hypergraph:
    """Column-net hypergraph of an n x n sparse matrix (net = row with its
    diagonal vertex as source, vertex = column); see fp_gen_params."""
    lib = _native.host()
    p = _native.GenParams(n, degree_mode, deg_lo, deg_hi, alpha, seed)
    net_ptr = np.zeros(n + 1, dtype=np.uint32)
    pins = ctypes.c_longlong(0)
    _native.check(lib.fp_gen_colnet_sizes(ctypes.byref(p), _ptr(net_ptr, _U32P), ctypes.byref(pins)),
                  "fp_gen_colnet_sizes")
    net_pins = np.zeros(pins.value, dtype=np.uint32)
    src = np.zeros(n, dtype=np.int32)
    _native.check(lib.fp_gen_colnet_pins(ctypes.byref(p), _ptr(net_ptr, _U32P), _ptr(net_pins, _U32P),
                                         _ptr(src, _I32P)), "fp_gen_colnet_pins")
    vwgt = None
    if weights is not None:
        vwgt = np.zeros(n, dtype=np.int32)
        _native.check(lib.fp_gen_weights(n, ctypes.c_ulonglong(seed), weights[0], weights[1],
                                         _ptr(vwgt, _I32P)), "fp_gen_weights")
    return Hypergraph.from_nets(n, net_ptr, net_pins, src, vwgt)


# The benchmark workloads, by what they are:
#   "c1" 10k x 10k, diagonal + 7 uniform off-diagonals per row (80,000 pins), K=16
#   "c2" 1M x 1M, total row degree ~ Zipf(2.1) on [2, 5000], K=256
#   "c3" 2M x 2M, diagonal + 12 uniform off-diagonals (26M pins), weights U{1..8}, K=1024
def workload(name: str, seed: int = 1, scale: float = 1.0) -> Hypergraph:
    if name == "c1":
        return colnet(max(2, int(10_000 * scale)), DEG_FIXED, 7, seed=seed)
    if name == "c2":
        return colnet(max(2, int(1_000_000 * scale)), DEG_ZIPF, 2, 5000, 2.1, seed=seed)
    if name == "c3":
        return colnet(max(2, int(2_000_000 * scale)), DEG_FIXED, 12, seed=seed, weights=(1, 8))
    raise ValueError(f"unknown workload {name!r}")


def read_patoh(path: str, src=None) -> Hypergraph:
    lib = _native.host()
    base, nov, noe, pins = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
    _native.check(lib.fp_read_patoh_header(path.encode(), ctypes.byref(base), ctypes.byref(nov),
                                           ctypes.byref(noe), ctypes.byref(pins)), "fp_read_patoh_header")
    cap_e = max(noe.value, 1)
    net_ptr = np.zeros(cap_e + 1, dtype=np.uint32)
    net_pins = np.zeros(max(pins.value, 1), dtype=np.uint32)
    rn, re_, rp = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
    _native.check(lib.fp_read_patoh(path.encode(), nov.value, cap_e, pins.value, ctypes.byref(rn),
                                    ctypes.byref(re_), ctypes.byref(rp), _ptr(net_ptr, _U32P),
                                    _ptr(net_pins, _U32P)), "fp_read_patoh")
    return Hypergraph.from_nets(rn.value, net_ptr[: re_.value + 1], net_pins[: rp.value], src)


def read_hypergraph(path: str, src=None, write_cache: bool = True) -> tuple[Hypergraph, bool]:
    """The reference reader's behaviour (fp_read_hypergraph): <path>.plbin
    when it exists, else the text (then the cache is written). Returns the
    hypergraph and whether it came from the cache."""
    lib = _native.host()
    bin_path = path + ".plbin"
    nov, noe, pins = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
    if os.path.exists(bin_path):
        _native.check(lib.fp_read_plbin_header(bin_path.encode(), ctypes.byref(nov), ctypes.byref(noe),
                                               ctypes.byref(pins)), "fp_read_plbin_header")
    else:
        base = ctypes.c_int()
        _native.check(lib.fp_read_patoh_header(path.encode(), ctypes.byref(base), ctypes.byref(nov),
                                               ctypes.byref(noe), ctypes.byref(pins)), "fp_read_patoh_header")
    cap_e = max(noe.value, 1)
    net_ptr = np.zeros(cap_e + 1, dtype=np.uint32)
    net_pins = np.zeros(max(pins.value, 1), dtype=np.uint32)
    rn, re_, rp, cached = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int()
    _native.check(lib.fp_read_hypergraph(path.encode(), int(write_cache), nov.value, cap_e, pins.value,
                                         ctypes.byref(rn), ctypes.byref(re_), ctypes.byref(rp), _ptr(net_ptr, _U32P),
                                         _ptr(net_pins, _U32P), ctypes.byref(cached)), "fp_read_hypergraph")
    return Hypergraph.from_nets(rn.value, net_ptr[: re_.value + 1], net_pins[: rp.value], src), bool(cached.value)


def write_plbin(hg: Hypergraph, path: str) -> None:
    _native.check(_native.host().fp_write_plbin(path.encode(), hg.nov, hg.noe, _ptr(hg.net_ptr, _U32P),
                                                _ptr(hg.net_pins, _U32P)), "fp_write_plbin")


def ws_path(hg_path: str) -> str:
    buf = ctypes.create_string_buffer(len(hg_path) + 8)
    _native.check(_native.host().fp_ws_path(hg_path.encode(), buf, len(buf)), "fp_ws_path")
    return buf.value.decode()


def read_ws(path: str, noe: int):
    """(edge weights, sources) of a .ws file."""
    w = np.zeros(noe, np.int32)
    src = np.zeros(noe, np.int32)
    _native.check(_native.host().fp_read_ws(path.encode(), noe, _ptr(w, _I32P), _ptr(src, _I32P)), "fp_read_ws")
    return w, src


def write_ws(path: str, src, edge_w=None) -> None:
    src = np.ascontiguousarray(src, dtype=np.int32)
    ew = None if edge_w is None else _ptr(np.ascontiguousarray(edge_w, dtype=np.int32), _I32P)
    _native.check(_native.host().fp_write_ws(path.encode(), len(src), ew, _ptr(src, _I32P)), "fp_write_ws")


def write_part_vector(path: str, part) -> None:
    part = np.ascontiguousarray(part, dtype=np.int32)
    _native.check(_native.host().fp_write_part_vector(path.encode(), len(part), _ptr(part, _I32P)),
                  "fp_write_part_vector")


def read_part_vector(path: str, nov: int) -> np.ndarray:
    part = np.zeros(nov, np.int32)
    _native.check(_native.host().fp_read_part_vector(path.encode(), nov, _ptr(part, _I32P)), "fp_read_part_vector")
    return part


def write_patoh(hg: Hypergraph, path: str, index_base: int = 0) -> None:
    lib = _native.host()
    _native.check(lib.fp_write_patoh(path.encode(), index_base, hg.nov, hg.noe, _ptr(hg.net_ptr, _U32P),
                                     _ptr(hg.net_pins, _U32P)), "fp_write_patoh")


def initial_partition(hg: Hypergraph, K: int, C: float, max_pw: int, min_pw: int, bands: bool = True,
                      ordering=None) -> np.ndarray:
    """Streaming LDG initial partition (fp_initial_partition)."""
    lib = _native.host()
    part = np.zeros(hg.nov, dtype=np.int32)
    ordp = None if ordering is None else _ptr(np.ascontiguousarray(ordering, dtype=np.int32), _I32P)
    vw = None if hg.vwgt is None else _ptr(hg.vwgt, _I32P)
    rc = lib.fp_initial_partition(0 if bands else 1, hg.nov, hg.noe, _ptr(hg.net_ptr, _U32P),
                                  _ptr(hg.vtx_ptr, _U32P), _ptr(hg.vtx_nets, _U32P), ordp, vw, K,
                                  ctypes.c_double(C), ctypes.c_longlong(max_pw), ctypes.c_longlong(min_pw),
                                  _ptr(part, _I32P))
    _native.check(rc, "fp_initial_partition")
    return part


INIT_LDG_BANDS, INIT_LDG, INIT_CN_LDG, INIT_MSG_GREEDY = 0, 1, 2, 3
# each objective's own initial stream in the reference
INIT_OF = {"TV": INIT_LDG_BANDS, "CN": INIT_CN_LDG, "MSV": INIT_LDG, "MSRV": INIT_LDG,
           "TM": INIT_MSG_GREEDY, "MSM": INIT_MSG_GREEDY}


def initial_partition_mode(hg: Hypergraph, mode: int, K: int, C: float, max_pw: int = 0, min_pw: int = 0,
                           ordering=None) -> np.ndarray:
    """Streaming initial partition of the given mode (fp_initial_partition,
    fp_initial_partition_msg): LDG2 with bands, band-free LDG, the CN stream or
    the TM/MSM message greedy."""
    lib = _native.host()
    part = np.zeros(hg.nov, dtype=np.int32)
    ordp = None if ordering is None else _ptr(np.ascontiguousarray(ordering, dtype=np.int32), _I32P)
    vw = None if hg.vwgt is None else _ptr(hg.vwgt, _I32P)
    if mode == INIT_MSG_GREEDY:
        rc = lib.fp_initial_partition_msg(hg.nov, hg.noe, _ptr(hg.net_ptr, _U32P), _ptr(hg.vtx_ptr, _U32P),
                                          _ptr(hg.vtx_nets, _U32P), _ptr(hg.src, _I32P), ordp, vw, K,
                                          ctypes.c_double(C), _ptr(part, _I32P))
        _native.check(rc, "fp_initial_partition_msg")
        return part
    rc = lib.fp_initial_partition(mode, hg.nov, hg.noe, _ptr(hg.net_ptr, _U32P), _ptr(hg.vtx_ptr, _U32P),
                                  _ptr(hg.vtx_nets, _U32P), ordp, vw, K, ctypes.c_double(C),
                                  ctypes.c_longlong(max_pw), ctypes.c_longlong(min_pw), _ptr(part, _I32P))
    _native.check(rc, "fp_initial_partition")
    return part


def initial_partition_for(metric: str, hg: Hypergraph, K: int, eps: float = 0.05, ordering=None) -> np.ndarray:
    """The initial partition the reference refines `metric` from, with the
    reference's constants at imbalance eps (C, max_pw, min_pw of the total
    vertex weight)."""
    C, max_pw, min_pw = ref_constants(hg.total_weight(), K, eps)
    return initial_partition_mode(hg, INIT_OF[metric], K, C, max_pw, min_pw, ordering)


def ref_constants(total_weight: int, K: int, eps: float = 0.05):
    """(C, max_pw, min_pw): the reference's LDG constants (TV/partition.cpp:95-96,159-160)."""
    lib = _native.host()
    return (lib.fp_ref_capacity(total_weight, K, ctypes.c_double(eps)),
            lib.fp_ref_max_pw(total_weight, K), lib.fp_ref_min_pw(total_weight, K))


def balance_cap(total_weight: int, K: int, eps: float) -> int:
    return int(_native.host().fp_balance_cap(total_weight, K, ctypes.c_double(eps)))
