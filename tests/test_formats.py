"""Readers and file formats of the host library against the reference's own
reader (read_patohlike, shared/hypergraph_rw.cpp:97-242, built unchanged into
oracle/_ref/ref_reader): the same CSR from the same text, including 1-based
files, blank lines and the noe correction; .plbin caches byte-identical to
the reference's and readable both ways; .ws and part-vector files in the
reference's layout."""
import os
import shutil

import numpy as np
import pytest

from fastpartitioner_amd import hypergraph as H
from fastpartitioner_amd._native import FastPartError
from oracle import oracle as O

needs_reader = pytest.mark.skipif(not O.ref_available("reader"), reason="oracle/_ref/ref_reader not built here")


def _write_text(path, header, lines):
    with open(path, "w") as f:
        f.write(header + "\n")
        for ln in lines:
            f.write(ln + "\n")


def _cases():
    hg = H.colnet(3000, H.DEG_ZIPF, 2, 200, 2.1, seed=4)
    rows = [" ".join(str(int(v)) for v in hg.net_pins[hg.net_ptr[n]:hg.net_ptr[n + 1]]) for n in range(hg.noe)]
    rows1 = [" ".join(str(int(v) + 1) for v in hg.net_pins[hg.net_ptr[n]:hg.net_ptr[n + 1]]) for n in range(hg.noe)]
    with_blanks = []
    for i, r in enumerate(rows):
        with_blanks.append(r)
        if i % 97 == 0:
            with_blanks.append("")
    return {
        "base0": (f"0 {hg.nov} {hg.noe} {hg.pins}", rows),
        "base1": (f"1 {hg.nov} {hg.noe} {hg.pins}", rows1),
        "blank_lines": (f"0 {hg.nov} {hg.noe} {hg.pins}", with_blanks),
        # the header claims more nets than the file holds: noe is corrected
        "noe_corrected": (f"0 {hg.nov} {hg.noe + 50} {hg.pins}", rows),
        # trailing spaces, tabs
        "spacing": (f"0 {hg.nov} {hg.noe} {hg.pins}  ", [r.replace(" ", " \t ") + "  " for r in rows]),
    }


@needs_reader
@pytest.mark.parametrize("name", sorted(_cases()))
def test_reader_matches_read_patohlike(tmp_path, name):
    header, lines = _cases()[name]
    path = str(tmp_path / f"{name}.hg")
    _write_text(path, header, lines)
    ref = O.run_reference_reader(path)   # also writes path.plbin, as the reference does
    assert ref["ok"]
    os.remove(path + ".plbin")
    hg, cached = H.read_hypergraph(path, write_cache=False)
    assert not cached
    assert hg.nov == ref["nov"] and hg.noe == ref["noe"] and hg.pins == ref["pins"]
    assert np.array_equal(hg.net_ptr, ref["net_ptr"]) and np.array_equal(hg.net_pins, ref["net_pins"])


@needs_reader
def test_pin_count_mismatch_fails_like_the_reference(tmp_path):
    header, lines = _cases()["base0"]
    nov, noe, pins = header.split()[1:]
    path = str(tmp_path / "bad.hg")
    _write_text(path, f"0 {nov} {noe} {int(pins) + 1}", lines)
    assert not O.run_reference_reader(path)["ok"]
    os.remove(path + ".plbin") if os.path.exists(path + ".plbin") else None
    with pytest.raises(FastPartError):
        H.read_hypergraph(path, write_cache=False)


@needs_reader
def test_plbin_cache_is_the_references(tmp_path):
    header, lines = _cases()["base1"]
    ref_dir, own_dir = tmp_path / "ref", tmp_path / "own"
    ref_dir.mkdir()
    own_dir.mkdir()
    ref_path, own_path = str(ref_dir / "g.hg"), str(own_dir / "g.hg")
    _write_text(ref_path, header, lines)
    shutil.copy(ref_path, own_path)
    first = O.run_reference_reader(ref_path)           # the reference writes its cache
    hg, cached = H.read_hypergraph(own_path)            # we write ours
    assert not cached
    with open(ref_path + ".plbin", "rb") as a, open(own_path + ".plbin", "rb") as b:
        assert a.read() == b.read()
    # each side reads the other's cache (the text is not read again)
    os.remove(ref_path)
    os.remove(own_path)
    shutil.copy(ref_path + ".plbin", own_path + ".plbin")
    again, cached = H.read_hypergraph(own_path)
    assert cached and np.array_equal(again.net_pins, first["net_pins"]) and np.array_equal(again.net_ptr,
                                                                                          first["net_ptr"])
    H.write_plbin(hg, ref_path + ".plbin")
    back = O.run_reference_reader(ref_path)
    assert back["ok"] and "Binary file detected" in back["stdout"]
    assert np.array_equal(back["net_pins"], hg.net_pins) and np.array_equal(back["net_ptr"], hg.net_ptr)


def test_ws_and_part_vector_files(tmp_path):
    hg = H.colnet(500, H.DEG_UNIFORM, 1, 5, seed=2)
    assert H.ws_path("/a/b/graph.mtx.hg") == "/a/b/graph.mtx.ws"
    assert H.ws_path("/a/b/graph") == "/a/b/graph.ws"
    ws = str(tmp_path / "g.ws")
    wts = np.arange(hg.noe, dtype=np.int32) % 10
    H.write_ws(ws, hg.src, wts)
    lines = open(ws).read().splitlines()
    assert len(lines) == hg.noe and lines[3] == f"{wts[3]} {hg.src[3]}"   # "weight source" per net
    w, src = H.read_ws(ws, hg.noe)
    assert np.array_equal(w, wts) and np.array_equal(src, hg.src)
    part = np.random.default_rng(1).integers(0, 37, hg.nov).astype(np.int32)
    pv = str(tmp_path / "g.hg.fastpart.37")
    H.write_part_vector(pv, part)
    assert open(pv).read() == "".join(f"{int(p)}\n" for p in part)
    assert np.array_equal(H.read_part_vector(pv, hg.nov), part)


def test_parallel_reader_on_a_large_file(tmp_path):
    hg = H.colnet(200_000, H.DEG_ZIPF, 2, 5000, 2.1, seed=9)
    path = str(tmp_path / "big.hg")
    H.write_patoh(hg, path, index_base=1)
    got, _ = H.read_hypergraph(path, write_cache=False)
    assert np.array_equal(got.net_ptr, hg.net_ptr) and np.array_equal(got.net_pins, hg.net_pins)
