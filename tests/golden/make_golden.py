"""Writes the golden fixtures of tests/golden/ by running the REFERENCE's own
CPU partitioner (oracle/_ref/ref_<m>, compiled from the reference sources by
oracle/Makefile) on small seeded instances. Only data is committed: the
instance (net->pins CSR and sources) and the states the reference reached.

Run from the repository root where the reference tree is available:
    make -C oracle ref && python tests/golden/make_golden.py
    make -C oracle cli && python tests/golden/make_golden.py bfs
"""
from __future__ import annotations

import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fastpartitioner_amd import hypergraph as H  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# (name, reference binary, K, passes, threads: 0 = partition(), 1 = partition_par
#  with one thread, instance). Names ending in "_perm" hand the reference a
#  seeded random visiting order (stored in the fixture as `ordering`).
CASES = [
    ("tv_serial", "tv", 8, 5, 0, "fixed5"),
    ("tv_par1", "tv", 8, 5, 1, "fixed5"),
    ("cn_serial", "cn", 8, 5, 0, "fixed5"),
    ("cn_par1", "cn", 8, 5, 1, "fixed5"),
    ("msv_serial", "mgv", 8, 5, 0, "fixed5"),
    ("msv_par1", "mgv", 8, 5, 1, "fixed5"),
    ("tm_serial", "tm", 64, 4, 0, "sparse"),
    ("tm_par1", "tm", 64, 4, 1, "sparse"),
    ("msm_serial", "msm", 256, 6, 0, "sparse2"),
    ("msm_par1", "msm", 256, 6, 1, "sparse2"),
    ("msrv_serial", "msrv", 4, 6, 0, "thin"),
    ("msrv_par1_perm", "msrv", 6, 6, 1, "mid"),
    ("tv_par1_perm", "tv", 8, 5, 1, "fixed5"),
]
PRINTED = ("TV", "CN", "TM", "MGM", "MGV", "MGAV")


def instance(kind):
    if kind == "fixed5":
        # 2,000 x 2,000 column-net matrix: diagonal + 5 uniform off-diagonals per row
        return H.colnet(2000, H.DEG_FIXED, 5, seed=7)
    if kind == "sparse":
        # 1,500 x 1,500: diagonal + 1..3 uniform off-diagonals (TM moves at K = 64)
        return H.colnet(1500, H.DEG_UNIFORM, 1, 3, seed=3)
    if kind == "thin":
        # 1,500 x 1,500: diagonal + 1..2 uniform off-diagonals (MSRV moves at K = 4)
        return H.colnet(1500, H.DEG_UNIFORM, 1, 2, seed=2)
    if kind == "mid":
        # 2,000 x 2,000: diagonal + 2..5 uniform off-diagonals
        return H.colnet(2000, H.DEG_UNIFORM, 2, 5, seed=2)
    # 3,000 x 3,000: diagonal + 1..3 uniform off-diagonals (MSM moves at K = 256)
    return H.colnet(3000, H.DEG_UNIFORM, 1, 3, seed=2)


def main(only=None):
    for name, m, K, passes, threads, kind in CASES:
        if only and name not in only:
            continue
        hg = instance(kind)
        if not O.ref_available(m):
            raise SystemExit(f"oracle/_ref/ref_{m} not built: run `make -C oracle ref`")
        ordering = None
        if name.endswith("_perm"):
            ordering = np.random.default_rng(K * 1000 + passes).permutation(hg.nov).astype(np.int32)
        r = O.run_reference(m, hg, K, passes, threads, ordering=ordering)
        printed = {}
        for key in PRINTED:
            mm = re.findall(rf"^After Refinement {key} : (-?\d+)$", r["stdout"], re.M)
            printed[key] = int(mm[-1])
        extra = {} if ordering is None else {"ordering": ordering}
        np.savez_compressed(
            os.path.join(OUT, f"{name}.npz"), **extra,
            nov=hg.nov, net_ptr=hg.net_ptr, net_pins=hg.net_pins, src=hg.src, K=K, passes=passes,
            threads=threads, init_part=r["init_part"], init_ends=r["init_ends"], init_parts=r["init_parts"],
            part=r["part"], ends=r["ends"], parts=r["parts"], cnt=r["cnt"],
            printed=np.array([printed[k] for k in PRINTED], dtype=np.int64))
        print(name, {k: printed[k] for k in PRINTED})


def write_bfs():
    """bfs.npz: the reference's BFS ordering (oracle/_ref/ref_bfs, built by
    `make -C oracle cli`) on an instance whose nets list their pins in a
    shuffled order, through both net views."""
    base = H.colnet(4000, H.DEG_UNIFORM, 1, 3, seed=26)
    rng = np.random.default_rng(26)
    pins = np.array(base.net_pins, dtype=np.uint32)
    for n in range(base.noe):
        rng.shuffle(pins[int(base.net_ptr[n]):int(base.net_ptr[n + 1])])
    hg = H.Hypergraph.from_nets(base.nov, base.net_ptr, pins, base.src)
    start = 1804289383 % hg.nov
    out = {}
    for view, key in (("ascending", "asc"), ("file", "file")):
        order, f1, f2 = O.run_reference_bfs(hg, start, view)
        out[f"{key}_ordering"] = order
        out[f"{key}_farthest"] = np.array([f1, f2], dtype=np.int32)
    np.savez_compressed(os.path.join(OUT, "bfs.npz"), nov=hg.nov, net_ptr=hg.net_ptr, net_pins=hg.net_pins,
                        src=hg.src, start=start, **out)
    print("bfs", {k: v.tolist() for k, v in out.items() if k.endswith("farthest")})


if __name__ == "__main__":
    if sys.argv[1:] == ["bfs"]:
        write_bfs()
    else:
        main(sys.argv[1:])
