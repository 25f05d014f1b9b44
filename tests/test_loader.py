"""Loader parity (CPU): the product's fast parser against the oracle's
restatement of graph_t::from_edges / deserialize, on the committed quirk
fixtures and on randomised text."""
import json
import os
import random

import numpy as np
import pytest

import nuc34
import oracle34
from conftest import GOLDEN

EXPECTED = json.load(open(os.path.join(GOLDEN, "expected.json")))


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_fixture_csr(name):
    exp = EXPECTED[name]
    path = os.path.join(GOLDEN, name)
    if "error" in exp:
        with pytest.raises(nuc34.Nuc34Error) as ei:
            nuc34.load_dataset(path)
        assert exp["error"] in str(ei.value)
        return
    g = nuc34.load_dataset(path)
    assert g.n == exp["n"]
    if name.endswith(".grh"):
        assert g.oriented
        O, E = oracle34.OracleGraph.load(path, orient=False).csr()
    else:
        assert not g.oriented
        O, E = np.array(exp["und_O"], np.uint32), np.array(exp["und_E"], np.uint32)
    assert np.array_equal(g.O, O) and np.array_equal(g.E, E)


def test_first_line_is_dropped():
    g = nuc34.load_dataset(os.path.join(GOLDEN, "first_line.edges"))
    # "1 2" sits on the first line only, so vertices 1 and 2 are not adjacent
    assert 2 not in g.E[g.O[1]:g.O[2]]


def test_mtx_ids_are_not_shifted():
    g = nuc34.load_dataset(os.path.join(GOLDEN, "onebased.mtx"))
    assert g.O[1] - g.O[0] == 0  # vertex 0 is isolated in a 1-based file
    assert g.n == 7


def _random_text(rng, lines):
    out = []
    for _ in range(rng.randrange(0, 3)):
        out.append(rng.choice(["# c", "%c", "#"]))
    out.append(rng.choice(["0 0", "5 5 5", "1 2"]))
    for _ in range(lines):
        s, d = rng.randrange(0, 30), rng.randrange(0, 30)
        form = rng.randrange(8)
        if form == 0:
            out.append(f"{s}\t{d}")
        elif form == 1:
            out.append(f"{s} {d} {rng.random():.3f}")
        elif form == 2:
            out.append(f"  {s}   {d}  ")
        elif form == 3:
            out.append(f"+{s} {d}\r")
        elif form == 4:
            out.append(f"{s} x")  # second token not a number: destination 0
        elif form == 5:
            out.append(f"{s},{d}")  # ',' ends the first number: destination 0
        else:
            out.append(f"{s} {d}")
    return "\n".join(out) + ("\n" if rng.random() < 0.8 else "")


@pytest.mark.parametrize("seed", range(40))
def test_random_text_matches_oracle(tmp_path, seed):
    rng = random.Random(seed)
    ext = rng.choice([".edges", ".txt"])
    p = tmp_path / f"g{seed}{ext}"
    p.write_text(_random_text(rng, rng.randrange(0, 60)), newline="")
    g = nuc34.load_dataset(str(p))
    O, E = oracle34.OracleGraph.load(str(p), orient=False).csr()
    assert np.array_equal(g.O, O) and np.array_equal(g.E, E)


@pytest.mark.parametrize("seed", range(6))
def test_host_orientation_matches_oracle(tmp_path, seed):
    from conftest import gen
    p = gen("rmat", 9, 8, 0.57, 0.19, 0.19, seed, out=tmp_path / "r.edges")
    h = nuc34.lib()
    import ctypes as C
    hg = C.c_void_p()
    err = C.create_string_buffer(256)
    assert h.nuc34_graph_load(p.encode(), C.byref(hg), err, 256) == 0
    assert h.nuc34_graph_orient_host(hg) == 0
    n, m, o = C.c_uint32(), C.c_uint64(), C.c_int()
    h.nuc34_graph_info(hg, C.byref(n), C.byref(m), C.byref(o))
    O = np.empty(n.value + 1, np.uint32)
    E = np.empty(m.value, np.uint32)
    h.nuc34_graph_copy(hg, O.ctypes.data_as(C.POINTER(C.c_uint32)), E.ctypes.data_as(C.POINTER(C.c_uint32)))
    h.nuc34_graph_free(hg)
    Oo, Eo = oracle34.OracleGraph.load(p, orient=True).csr()
    assert o.value == 1 and np.array_equal(O, Oo) and np.array_equal(E, Eo)


def test_pairs_match_text(tmp_path):
    rng = np.random.default_rng(3)
    s = rng.integers(0, 50, 300, dtype=np.uint32)
    d = rng.integers(0, 50, 300, dtype=np.uint32)
    p = tmp_path / "x.edges"
    p.write_text("0 0\n" + "".join(f"{a} {b}\n" for a, b in zip(s, d)))
    a = nuc34.graph_from_pairs(s, d)
    b = nuc34.load_dataset(str(p))
    assert np.array_equal(a.O, b.O) and np.array_equal(a.E, b.E)


@pytest.mark.parametrize("ext", [".edges", ".mtx"])
def test_large_text_slices_match_oracle(tmp_path, ext):
    """Texts of several MiB are scanned in newline-aligned slices by parallel
    threads; ids span more than 4096 vertices so the CSR build uses several
    source buckets. Result must equal the serial oracle parse."""
    rng = random.Random(11)
    out = ["% c", "9000 9000 0"] if ext == ".mtx" else ["# c", "1 2"]
    forms = 0
    for i in range(400_000):
        s, d = rng.randrange(0, 9000), rng.randrange(0, 9000)
        if i % 97 == 0:
            s = d = rng.randrange(0, 9000)       # self loop
        if i % 1013 == 0:
            out.append("# not a comment here")    # no leading number: skipped
            forms += 1
        if i % 5 == 0:
            out.append(f"{s}\t{d} 0.5\r")
        elif i % 7 == 0:
            out.append(f"  {d}  {s}")             # reversed duplicate direction
        else:
            out.append(f"{s} {d}")
    p = tmp_path / f"big{ext}"
    p.write_text("\n".join(out), newline="")   # no trailing newline
    assert p.stat().st_size > 4 << 20 and forms > 0
    g = nuc34.load_dataset(str(p))
    O, E = oracle34.OracleGraph.load(str(p), orient=False).csr()
    assert np.array_equal(g.O, O) and np.array_equal(g.E, E)


def test_large_pairs_match_numpy():
    """More than a million pairs are split across threads; the CSR equals a
    numpy sort/unique of both directions with self loops dropped."""
    rng = np.random.default_rng(5)
    cnt = 3_000_000
    s = (rng.zipf(1.6, cnt) % 200_000).astype(np.uint32)
    d = rng.integers(0, 200_000, cnt, dtype=np.uint32)
    g = nuc34.graph_from_pairs(s, d)
    keep = s != d
    a = np.concatenate([s[keep], d[keep]]).astype(np.uint64)
    b = np.concatenate([d[keep], s[keep]]).astype(np.uint64)
    keys = np.unique((a << np.uint64(32)) | b)
    n = int(max(s[keep].max(), d[keep].max())) + 1
    src = (keys >> np.uint64(32)).astype(np.int64)
    O = np.zeros(n + 1, np.uint64)
    np.add.at(O, src + 1, 1)
    O = np.cumsum(O).astype(np.uint32)
    assert g.n == n
    assert np.array_equal(g.O, O)
    assert np.array_equal(g.E, (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))
