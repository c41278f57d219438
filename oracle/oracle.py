"""ctypes binding of liboracle.so (fp_oracle.c) and of the reference harness
binaries in oracle/_ref. TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.
"""
from __future__ import annotations

import ctypes
import os
import re
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_U32P = ctypes.POINTER(ctypes.c_uint32)
_I32P = ctypes.POINTER(ctypes.c_int32)
_I64P = ctypes.POINTER(ctypes.c_int64)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle` or __graft_entry__.build()")
        _lib = ctypes.CDLL(LIB)
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _wts(hg):
    """Hands hg's vertex weights (None = unit) to the oracle's refinements and
    initial partitions, which change part sizes by the moved vertex's weight."""
    lib().or_set_vertex_weights(_p(hg.vwgt, _I32P))


def _ord(hg, ordering):
    if ordering is None:
        ordering = np.arange(hg.nov, dtype=np.int32)
    return np.ascontiguousarray(ordering, dtype=np.int32)


class State:
    """Connectivity state in the reference's layout (TV/partition.cpp:150-158)."""

    def __init__(self, hg, K, sizes_dtype=np.int32):
        self.K = K
        self.part = np.full(hg.nov, -1, dtype=np.int32)
        self.ep_ends = np.zeros(hg.noe, dtype=np.int32)
        self.ep_parts = np.full(hg.pins, -1, dtype=np.int32)
        self.ep_cnt = np.zeros(hg.pins, dtype=np.int32)
        self.sizes = np.zeros(K, dtype=sizes_dtype)

    def copy(self):
        s = State.__new__(State)
        s.K = self.K
        for k in ("part", "ep_ends", "ep_parts", "ep_cnt", "sizes", "sv", "rv", "pc", "msg"):
            if hasattr(self, k):
                setattr(s, k, getattr(self, k).copy())
        return s


def build_state(hg, K, part, ordering=None):
    st = State(hg, K, np.int64)
    st.part[:] = part
    rc = lib().or_build_state(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                              _p(hg.vtx_nets, _U32P), _p(_ord(hg, ordering), _I32P), K, _p(st.part, _I32P),
                              _p(hg.vwgt, _I32P), _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P),
                              _p(st.ep_cnt, _I32P), _p(st.sizes, _I64P))
    if rc:
        raise ValueError(f"or_build_state rc={rc}")
    return st


METRIC_NAMES = ("TV", "CN", "TM", "MSM", "MSV", "MSRV")


def metrics(hg, K, part):
    out = np.zeros(6, dtype=np.int64)
    lam = np.zeros(hg.noe, dtype=np.int32)
    vec = {k: np.zeros(K, dtype=np.int64) for k in ("sv", "rv", "sm", "rm", "sv_msv")}
    rc = lib().or_metrics(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.net_pins, _U32P), _p(hg.src, _I32P),
                          K, _p(np.ascontiguousarray(part, dtype=np.int32), _I32P), _p(out, _I64P),
                          _p(lam, _I32P), _p(vec["sv"], _I64P), _p(vec["rv"], _I64P), _p(vec["sm"], _I64P),
                          _p(vec["rm"], _I64P), _p(vec["sv_msv"], _I64P))
    if rc:
        raise ValueError(f"or_metrics rc={rc}")
    res = dict(zip(METRIC_NAMES, (int(x) for x in out)))
    res["lam"] = lam
    res.update(vec)
    return res


def tv_ldg(hg, K, C, max_pw, min_pw, par_skip=False, ordering=None):
    _wts(hg)
    st = State(hg, K)
    marker = np.zeros(K, dtype=np.int32)
    rc = lib().or_tv_ldg(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P), _p(hg.vtx_nets, _U32P),
                         _p(_ord(hg, ordering), _I32P), K, ctypes.c_double(C), max_pw, min_pw, int(par_skip),
                         _p(st.part, _I32P), _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P),
                         _p(st.sizes, _I32P), _p(marker, _I32P))
    if rc:
        raise ValueError(f"or_tv_ldg rc={rc}")
    return st, marker


ROUND_LDG_MODES = {"ldg_bands": 0, "ldg": 1, "cn_ldg": 2}


def round_ldg(hg, K, mode, C, max_pw, R, cap, ordering=None):
    """Buffered round stream of fp_oracle.c's or_round_ldg (the engine's
    device initial partition); returns (State with int64 sizes, sub-rounds)."""
    _wts(hg)
    st = State(hg, K, np.int64)
    sub = ctypes.c_longlong()
    rc = lib().or_round_ldg(ROUND_LDG_MODES[mode], hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                            _p(hg.vtx_nets, _U32P), _p(_ord(hg, ordering), _I32P), K, ctypes.c_double(C),
                            ctypes.c_longlong(max_pw), R, ctypes.c_longlong(cap), _p(st.part, _I32P),
                            _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P),
                            _p(st.sizes, _I64P), ctypes.byref(sub))
    if rc:
        raise ValueError(f"or_round_ldg rc={rc}")
    return st, sub.value


def tv_refine_serial(hg, st, marker, C, max_pw, min_pw, passes, ordering=None):
    _wts(hg)
    rc = lib().or_tv_refine_serial(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                                   _p(hg.vtx_nets, _U32P), _p(_ord(hg, ordering), _I32P), st.K,
                                   ctypes.c_double(C), max_pw, min_pw, passes, _p(st.part, _I32P),
                                   _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P),
                                   _p(st.sizes, _I32P), _p(marker, _I32P))
    if rc:
        raise ValueError(f"or_tv_refine_serial rc={rc}")
    return st


def tv_refine_par(hg, st, C, max_pw, min_pw, passes, R, chunk=8, ordering=None, gains=None):
    _wts(hg)
    assert st.sizes.dtype == np.int32
    rc = lib().or_tv_refine_par(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                                _p(hg.vtx_nets, _U32P), _p(_ord(hg, ordering), _I32P), st.K, ctypes.c_double(C),
                                max_pw, min_pw, passes, R, chunk, _p(st.part, _I32P), _p(st.ep_ends, _I32P),
                                _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P), _p(st.sizes, _I32P),
                                _p(gains, _I32P))
    if rc:
        raise ValueError(f"or_tv_refine_par rc={rc}")
    return st


def tv_refine_bal(hg, st, C, max_pw, min_pw, cap, passes, R, ordering=None, gains=None):
    _wts(hg)
    assert st.sizes.dtype == np.int64
    L = ctypes.c_longlong
    rc = lib().or_tv_refine_bal(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                                _p(hg.vtx_nets, _U32P), _p(_ord(hg, ordering), _I32P), _p(hg.vwgt, _I32P),
                                st.K, ctypes.c_double(C), L(max_pw), L(min_pw), L(cap), passes, R,
                                _p(st.part, _I32P), _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P),
                                _p(st.ep_cnt, _I32P), _p(st.sizes, _I64P), _p(gains, _I32P))
    if rc:
        raise ValueError(f"or_tv_refine_bal rc={rc}")
    return st


def apply_move(hg, st, v, p, q):
    """The reference's list edits of one move (TV/partition.cpp:387-436)."""
    lib().or_apply_move(int(v), int(p), int(q), _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                        _p(hg.vtx_nets, _U32P), _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P),
                        _p(st.ep_cnt, _I32P))


def plant(code: int):
    """Test hook: plant a known mistake in the balanced admission (0 = none)."""
    lib().or_plant(int(code))


def cn_ldg(hg, K, C, ordering=None):
    _wts(hg)
    st = State(hg, K)
    rc = lib().or_cn_ldg(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P), _p(hg.vtx_nets, _U32P),
                         _p(_ord(hg, ordering), _I32P), K, ctypes.c_double(C), _p(st.part, _I32P),
                         _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P),
                         _p(st.sizes, _I32P))
    if rc:
        raise ValueError(f"or_cn_ldg rc={rc}")
    return st


def cn_refine(hg, st, passes, serial):
    _wts(hg)
    f = lib().or_cn_refine_serial if serial else lib().or_cn_refine_par
    rc = f(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.net_pins, _U32P), _p(hg.vtx_ptr, _U32P),
           _p(hg.vtx_nets, _U32P), st.K, passes, _p(st.part, _I32P), _p(st.ep_ends, _I32P),
           _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P), _p(st.sizes, _I32P))
    if rc:
        raise ValueError(f"or_cn_refine rc={rc}")
    return st


def msv_ldg(hg, K, C, ordering=None):
    """MGV initial partition; st.sv holds the send volumes it accumulated."""
    _wts(hg)
    st = State(hg, K)
    st.sv = np.zeros(K, dtype=np.int32)
    rc = lib().or_msv_ldg(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P), _p(hg.vtx_nets, _U32P),
                          _p(hg.src, _I32P), _p(_ord(hg, ordering), _I32P), K, ctypes.c_double(C),
                          _p(st.part, _I32P), _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P),
                          _p(st.sizes, _I32P), _p(st.sv, _I32P))
    if rc:
        raise ValueError(f"or_msv_ldg rc={rc}")
    return st


def msv_refine_serial(hg, st, C, passes, ordering=None):
    _wts(hg)
    rc = lib().or_msv_refine_serial(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                                    _p(hg.vtx_nets, _U32P), _p(hg.src, _I32P), _p(_ord(hg, ordering), _I32P), st.K,
                                    ctypes.c_double(C), passes, _p(st.part, _I32P), _p(st.ep_ends, _I32P),
                                    _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P), _p(st.sizes, _I32P),
                                    _p(st.sv, _I32P))
    if rc:
        raise ValueError(f"or_msv_refine_serial rc={rc}")
    return st


def msv_refine_par(hg, st, C, passes, R):
    """Returns the number of applied moves."""
    _wts(hg)
    assert st.sizes.dtype == np.int32 and st.sv.dtype == np.int32
    moves = ctypes.c_longlong(0)
    rc = lib().or_msv_refine_par(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                                 _p(hg.vtx_nets, _U32P), _p(hg.src, _I32P), st.K, ctypes.c_double(C), passes, R,
                                 _p(st.part, _I32P), _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P),
                                 _p(st.ep_cnt, _I32P), _p(st.sizes, _I32P), _p(st.sv, _I32P), ctypes.byref(moves))
    if rc:
        raise ValueError(f"or_msv_refine_par rc={rc}")
    return moves.value


def vol_move_delta(hg, st, v, d, recv):
    """losses - gains per part of the reference's evaluation of v -> d."""
    _wts(hg)
    out = np.zeros(st.K, dtype=np.int32)
    rc = lib().or_vol_move_delta(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                                 _p(hg.vtx_nets, _U32P), _p(hg.src, _I32P), st.K, _p(st.part, _I32P),
                                 _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P), int(recv),
                                 int(v), int(d), _p(out, _I32P))
    if rc:
        raise ValueError(f"or_vol_move_delta rc={rc}")
    return out


def tm_init(hg, K, C, ordering=None):
    """TM greedy initial partition; st.pc (K*K) and st.msg (K) hold the part connections."""
    _wts(hg)
    st = State(hg, K)
    st.pc = np.zeros(K * K, dtype=np.int32)
    st.msg = np.zeros(K, dtype=np.int32)
    rc = lib().or_tm_init(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P), _p(hg.vtx_nets, _U32P),
                          _p(hg.src, _I32P), _p(_ord(hg, ordering), _I32P), K, ctypes.c_double(C),
                          _p(st.part, _I32P), _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P),
                          _p(st.sizes, _I32P), _p(st.pc, _I32P), _p(st.msg, _I32P))
    if rc:
        raise ValueError(f"or_tm_init rc={rc}")
    return st


def msm_init(hg, K, C, ordering=None):
    """MSM greedy initial partition (pc diagonal -1)."""
    _wts(hg)
    st = State(hg, K)
    st.pc = np.zeros(K * K, dtype=np.int32)
    st.msg = np.zeros(K, dtype=np.int32)
    rc = lib().or_msm_init(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P), _p(hg.vtx_nets, _U32P),
                           _p(hg.src, _I32P), _p(_ord(hg, ordering), _I32P), K, ctypes.c_double(C),
                           _p(st.part, _I32P), _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P),
                           _p(st.sizes, _I32P), _p(st.pc, _I32P), _p(st.msg, _I32P))
    if rc:
        raise ValueError(f"or_msm_init rc={rc}")
    return st


def msm_refine_serial(hg, st, passes):
    _wts(hg)
    rc = lib().or_msm_refine_serial(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                                    _p(hg.vtx_nets, _U32P), _p(hg.src, _I32P), st.K, passes, *_conn_args(hg, st))
    if rc:
        raise ValueError(f"or_msm_refine_serial rc={rc}")
    return st


def msm_refine_par(hg, st, passes, R):
    _wts(hg)
    moves = ctypes.c_longlong(0)
    rc = lib().or_msm_refine_par(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                                 _p(hg.vtx_nets, _U32P), _p(hg.src, _I32P), st.K, passes, R, *_conn_args(hg, st),
                                 ctypes.byref(moves))
    if rc:
        raise ValueError(f"or_msm_refine_par rc={rc}")
    return moves.value


def _conn_args(hg, st):
    return (_p(st.part, _I32P), _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P),
            _p(st.sizes, _I32P), _p(st.pc, _I32P), _p(st.msg, _I32P))


def tm_refine_serial(hg, st, C, passes):
    _wts(hg)
    rc = lib().or_tm_refine_serial(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                                   _p(hg.vtx_nets, _U32P), _p(hg.src, _I32P), st.K, ctypes.c_double(C), passes,
                                   *_conn_args(hg, st))
    if rc:
        raise ValueError(f"or_tm_refine_serial rc={rc}")
    return st


def tm_refine_par(hg, st, C, passes, R):
    _wts(hg)
    moves = ctypes.c_longlong(0)
    rc = lib().or_tm_refine_par(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                                _p(hg.vtx_nets, _U32P), _p(hg.src, _I32P), st.K, ctypes.c_double(C), passes, R,
                                *_conn_args(hg, st), ctypes.byref(moves))
    if rc:
        raise ValueError(f"or_tm_refine_par rc={rc}")
    return moves.value


def msrv_ldg(hg, K, C, ordering=None):
    """MSRV initial partition; st.sv / st.rv hold the send / receive volumes."""
    _wts(hg)
    st = State(hg, K)
    st.sv = np.zeros(K, dtype=np.int32)
    st.rv = np.zeros(K, dtype=np.int32)
    rc = lib().or_msrv_ldg(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P), _p(hg.vtx_nets, _U32P),
                           _p(hg.src, _I32P), _p(_ord(hg, ordering), _I32P), K, ctypes.c_double(C),
                           _p(st.part, _I32P), _p(st.ep_ends, _I32P), _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P),
                           _p(st.sizes, _I32P), _p(st.sv, _I32P), _p(st.rv, _I32P))
    if rc:
        raise ValueError(f"or_msrv_ldg rc={rc}")
    return st


def msrv_refine_par(hg, st, C, passes, R, ordering=None):
    """Returns the number of applied moves."""
    _wts(hg)
    assert st.sizes.dtype == np.int32
    moves = ctypes.c_longlong(0)
    rc = lib().or_msrv_refine_par(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                                  _p(hg.vtx_nets, _U32P), _p(hg.src, _I32P), _p(_ord(hg, ordering), _I32P), st.K,
                                  ctypes.c_double(C), passes, R, _p(st.part, _I32P), _p(st.ep_ends, _I32P),
                                  _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P), _p(st.sizes, _I32P),
                                  _p(st.sv, _I32P), _p(st.rv, _I32P), ctypes.byref(moves))
    if rc:
        raise ValueError(f"or_msrv_refine_par rc={rc}")
    return moves.value


def msrv_refine_serial(hg, st, C, passes, ordering=None):
    _wts(hg)
    rc = lib().or_msrv_refine_serial(hg.nov, hg.noe, _p(hg.net_ptr, _U32P), _p(hg.vtx_ptr, _U32P),
                                     _p(hg.vtx_nets, _U32P), _p(hg.src, _I32P), _p(_ord(hg, ordering), _I32P), st.K,
                                     ctypes.c_double(C), passes, _p(st.part, _I32P), _p(st.ep_ends, _I32P),
                                     _p(st.ep_parts, _I32P), _p(st.ep_cnt, _I32P), _p(st.sizes, _I32P),
                                     _p(st.sv, _I32P), _p(st.rv, _I32P))
    if rc:
        raise ValueError(f"or_msrv_refine_serial rc={rc}")
    return st


# ---- reference harness (oracle/_ref/ref_<m>) -----------------------------

def ref_available(metric: str, binary: str = "ref") -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"{binary}_{metric}"))


def run_reference(metric: str, hg, K: int, passes: int, threads: int, timeout: float = 600.0,
                  env_extra=None, binary: str = "ref", ordering=None):
    """Run the reference's own partition()/partition_par() on `hg` and return
    the states it reached: dict with init_part/init_ends/init_parts and
    part/ends/parts/cnt (after refinement), the wall time and the
    'Refinement took' seconds it printed.

    binary="dropin" with threads=-1 runs the reference's partition_gpu(),
    whose device entry point is served by libfastpart_legacy_<metric>.so.
    ordering: a vertex permutation handed to the reference as its visiting
    order (default: the identity, as the reference's drivers build it)."""
    exe = os.path.join(REF_DIR, f"{binary}_{metric}")
    with tempfile.TemporaryDirectory() as d:
        inst = os.path.join(d, "g.fphg")
        out = os.path.join(d, "g.fprs")
        hg.write_fphg(inst)
        cmd = [exe, inst, str(K), str(passes), str(threads), out]
        if ordering is not None:
            opath = os.path.join(d, "g.ord")
            np.ascontiguousarray(ordering, dtype="<i4").tofile(opath)
            cmd.append(opath)
        env = dict(os.environ)
        if env_extra:
            env.update(env_extra)
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
        if p.returncode != 0:
            raise RuntimeError(f"{exe} failed rc={p.returncode}: {p.stderr[-2000:]}")
        with open(out, "rb") as f:
            raw = f.read()
    assert raw[:4] == b"FPRS"
    nov, noe, K_ = struct.unpack_from("<iii", raw, 4)
    pins, wall = struct.unpack_from("<qd", raw, 16)
    off = 32

    def take(n):
        nonlocal off
        a = np.frombuffer(raw, dtype="<i4", count=n, offset=off).copy()
        off += 4 * n
        return a

    res = {}
    for tag in ("init_", ""):
        res[tag + "part"] = take(nov)
        res[tag + "ends"] = take(noe)
        res[tag + "parts"] = take(pins)
    res["cnt"] = take(pins)
    res["wall"] = wall
    m = re.findall(r"Refinement took : ([0-9.eE+-]+) seconds", p.stdout)
    res["refine_seconds"] = float(m[-1]) if m else None
    res["stdout"] = p.stdout
    return res


def run_reference_reader(path: str, timeout: float = 300.0):
    """The reference's read_patohlike (oracle/_ref/ref_reader) on `path`: dict
    with ok, nov, noe, pins, net_ptr, net_pins, vertex_appearence, stdout.
    Like the reference, it reads <path>.plbin when that exists and otherwise
    writes it after reading the text."""
    exe = os.path.join(REF_DIR, "ref_reader")
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "r.bin")
        p = subprocess.run([exe, path, out], capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"{exe} failed rc={p.returncode}: {p.stderr[-2000:]}")
        with open(out, "rb") as f:
            raw = f.read()
    ok, nov, noe = struct.unpack_from("<iII", raw, 0)
    pins = struct.unpack_from("<q", raw, 12)[0]
    res = {"ok": bool(ok), "nov": nov, "noe": noe, "pins": pins, "stdout": p.stdout}
    if ok:
        off = 20
        res["net_ptr"] = np.frombuffer(raw, dtype="<u4", count=noe + 1, offset=off).copy()
        off += 4 * (noe + 1)
        res["net_pins"] = np.frombuffer(raw, dtype="<u4", count=pins, offset=off).copy()
        off += 4 * pins
        res["vertex_appearence"] = np.frombuffer(raw, dtype="<u4", count=nov, offset=off).copy()
    return res


def bfs_net_view(hg, view: str):
    """The net -> pins view a BFS expands: "ascending" = each net's pins in
    ascending vertex order (createInverse, shared/utils.cpp:32-56, which
    TV/MGV/TM drive the BFS with), "file" = the reader's pin order (CN/cn.cpp
    drives it with col_ptr/row_ind)."""
    if view == "file":
        return np.ascontiguousarray(hg.net_pins, dtype=np.uint32)
    if view != "ascending":
        raise ValueError(view)
    out = np.array(hg.net_pins, dtype=np.uint32)
    for n in range(hg.noe):
        s, e = int(hg.net_ptr[n]), int(hg.net_ptr[n + 1])
        out[s:e].sort()
    return out


def bfs_ordering(hg, start: int, view: str = "ascending"):
    """or_bfs_ordering: (ordering, farthest_first, farthest_all)."""
    pins = bfs_net_view(hg, view)
    ordering = np.zeros(hg.nov, dtype=np.int32)
    f1, f2 = ctypes.c_int(), ctypes.c_int()
    rc = lib().or_bfs_ordering(hg.nov, hg.noe, _p(hg.vtx_ptr, _U32P), _p(hg.vtx_nets, _U32P), _p(hg.net_ptr, _U32P),
                               _p(pins, _U32P), int(start), _p(ordering, _I32P), ctypes.byref(f1), ctypes.byref(f2))
    if rc:
        raise ValueError(f"or_bfs_ordering rc={rc}")
    return ordering, f1.value, f2.value


def run_reference_bfs(hg, start: int, view: str = "ascending", threads: int = 1, timeout: float = 300.0):
    """The reference's BFS ordering (oracle/_ref/ref_bfs): (ordering with -1
    for unreached vertices, farthest_first, farthest_all)."""
    exe = os.path.join(REF_DIR, "ref_bfs")
    pins = bfs_net_view(hg, view)
    with tempfile.TemporaryDirectory() as d:
        src, out = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(src, "wb") as f:
            f.write(struct.pack("<IIq", hg.nov, hg.noe, hg.pins))
            for a in (hg.vtx_ptr, hg.vtx_nets, hg.net_ptr, pins):
                f.write(np.ascontiguousarray(a, dtype="<u4").tobytes())
            f.write(struct.pack("<ii", int(start), int(threads)))
        p = subprocess.run([exe, src, out], capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"{exe} failed rc={p.returncode}: {p.stderr[-2000:]}")
        raw = open(out, "rb").read()
    f1, f2 = struct.unpack_from("<ii", raw, 0)
    return np.frombuffer(raw, dtype="<i4", count=hg.nov, offset=8).copy(), f1, f2
