"""Test helper: the oracle's refinement of a state for any objective and
schedule (oracle/fp_oracle.c through oracle/oracle.py), with the objective's
bookkeeping (send / receive volumes, part connections) recounted from the
state when the state does not carry it."""
import numpy as np

from fastpartitioner_amd.engine import SCHED_BAL
from oracle import oracle as O


def list_entries(hg, st):
    """(net, part, count) of every list entry, in list order."""
    lam = st.ep_ends.astype(np.int64)
    starts = hg.net_ptr[:-1].astype(np.int64)
    first = np.repeat(np.cumsum(lam) - lam, lam)
    idx = np.repeat(starts, lam) + (np.arange(int(lam.sum())) - first)
    return np.repeat(np.arange(hg.noe), lam), st.ep_parts[idx].astype(np.int64), st.ep_cnt[idx]


def with_connections(hg, st, msm, K):
    """Sets st.pc / st.msg, the part-connection matrix and message counts,
    recounted from the lists (MSM keeps -1 on the diagonal,
    MSM/partition.cpp:101)."""
    n, b, _ = list_entries(hg, st)
    a = st.part[hg.src].astype(np.int64)[n]
    off = b != a
    pc = np.bincount(a[off] * K + b[off], minlength=K * K).astype(np.int32)
    st.msg = (pc.reshape(K, K) > 0).sum(axis=1).astype(np.int32)
    if msm:
        pc.reshape(K, K)[np.arange(K), np.arange(K)] = -1
    st.pc = pc
    return st


def refine(hg, K, metric, sched, st, passes, ordering=None):
    """The oracle's refinement of a copy of st under sched (an engine
    Schedule, None for CN). Returns (state, applied moves or None)."""
    st = st.copy()
    if metric == "CN":
        st.sizes = st.sizes.astype(np.int32)
        O.cn_refine(hg, st, passes, serial=False)
        return st, None
    C, max_pw, min_pw = sched.ldg_C, int(sched.max_pw), int(sched.min_pw)
    if metric == "TV" and sched.kind == SCHED_BAL:
        st.sizes = st.sizes.astype(np.int64)
        O.tv_refine_bal(hg, st, C, sched.cap, 0, sched.cap, passes, sched.round, ordering=ordering)
        return st, None
    st.sizes = st.sizes.astype(np.int32)
    if metric == "TV":
        O.tv_refine_par(hg, st, C, max_pw, min_pw, passes, sched.round, ordering=ordering)
        return st, None
    if metric in ("MSV", "MSRV"):
        if not hasattr(st, "sv"):
            m = O.metrics(hg, K, st.part)
            st.sv = (m["sv_msv"] if metric == "MSV" else m["sv"]).astype(np.int32)
            if metric == "MSRV":
                st.rv = m["rv"].astype(np.int32)
        if metric == "MSV":
            return st, O.msv_refine_par(hg, st, C, passes, sched.round)
        return st, O.msrv_refine_par(hg, st, C, passes, sched.round, ordering=ordering)
    if not hasattr(st, "pc"):
        with_connections(hg, st, metric == "MSM", K)
    if metric == "TM":
        return st, O.tm_refine_par(hg, st, C, passes, sched.round)
    return st, O.msm_refine_par(hg, st, passes, sched.round)


def assert_same_state(got, st):
    for f in ("part", "ep_ends", "ep_parts", "ep_cnt"):
        assert np.array_equal(getattr(got, f), getattr(st, f)), f
    assert np.array_equal(got.sizes, st.sizes.astype(np.int64)), "sizes"
