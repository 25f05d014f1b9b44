"""Pins the oracle (CPU): closed forms, an independent brute force, the
committed golden vectors, and the reference's own host path compiled from its
headers (oracle/_ref/ref34) where that binary has been built."""
import itertools
import json
import math
import os
import subprocess

import numpy as np
import pytest

import brute
import oracle34
from conftest import GOLDEN, gen

EXPECTED = json.load(open(os.path.join(GOLDEN, "expected.json")))


def run_oracle(path, mode="host", parts=2):
    return oracle34.OracleGraph.load(path, orient=True).run(mode, parts)


def write_pairs(path, pairs):
    with open(path, "w") as f:
        f.write("0 0\n")
        for a, b in pairs:
            f.write(f"{a} {b}\n")
    return str(path)


@pytest.mark.parametrize("mode", ["host", "ondemand"])
@pytest.mark.parametrize("n", [3, 4, 5, 7, 9])
def test_clique_closed_form(tmp_path, n, mode):
    p = write_pairs(tmp_path / "k.edges", itertools.combinations(range(1, n + 1), 2))
    r = run_oracle(p, mode)
    assert r["T"] == math.comb(n, 3)
    assert r["Q"] == math.comb(n, 4)
    assert (r["support"] == n - 3).all()
    assert (r["k"] == n - 3).all()
    assert r["K_max"] == (n - 3 if n >= 4 else 0)


@pytest.mark.parametrize("mode", ["host", "ondemand"])
def test_disjoint_cliques(tmp_path, mode):
    sizes = [4, 6, 9]
    pairs, base, owner = [], 1, {}
    for s in sizes:
        vs = list(range(base, base + s))
        pairs += list(itertools.combinations(vs, 2))
        base += s
    p = write_pairs(tmp_path / "d.edges", pairs)
    r = run_oracle(p, mode)
    assert sorted(set(r["k"].tolist())) == [s - 3 for s in sizes]
    assert r["Q"] == sum(math.comb(s, 4) for s in sizes)
    assert r["K_max"] == max(sizes) - 3


def test_triangle_without_k4_gets_zero(tmp_path):
    # a K5 plus a pendant triangle hanging off one of its vertices
    pairs = list(itertools.combinations(range(1, 6), 2)) + [(5, 6), (6, 7), (5, 7)]
    r = run_oracle(write_pairs(tmp_path / "p.edges", pairs))
    assert r["T"] == 11 and r["Q"] == 5
    assert sorted(r["k"].tolist()) == [0] + [2] * 10


def _random_pairs(seed, n, p):
    rng = np.random.default_rng(seed)
    return [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]


@pytest.mark.parametrize("mode", ["host", "ondemand"])
@pytest.mark.parametrize("seed,n,p", [(1, 20, 0.5), (2, 30, 0.3), (3, 45, 0.25), (4, 60, 0.15), (5, 25, 0.7)])
def test_brute_force(tmp_path, seed, n, p, mode):
    pairs = _random_pairs(seed, n, p)
    path = write_pairs(tmp_path / "r.edges", pairs)
    r = run_oracle(path, mode, parts=3)
    nn, nbrs = brute.undirected_from_pairs(pairs)
    b = brute.decompose(nn, nbrs)
    assert r["T"] == b["T"] and r["Q"] == b["Q"] and r["K_max"] == b["K_max"]
    assert [tuple(t) for t in r["tris"].tolist()] == [tuple(t) for t in b["tris"]]
    assert r["support"].tolist() == b["support"]
    assert r["k"].tolist() == b["k"]


def test_brute_force_catches_planted_mistakes(tmp_path):
    """The comparisons above would notice a dropped K4, a swapped (a,b) or an
    off-by-one k: plant each in a copy of the oracle's output."""
    pairs = _random_pairs(9, 30, 0.4)
    r = run_oracle(write_pairs(tmp_path / "m.edges", pairs))
    nn, nbrs = brute.undirected_from_pairs(pairs)
    b = brute.decompose(nn, nbrs)
    tris = [tuple(t) for t in r["tris"].tolist()]
    swapped = [(t[1], t[0], t[2]) for t in tris]
    assert swapped != [tuple(t) for t in b["tris"]]
    k = r["k"].copy()
    k[len(k) // 2] += 1
    assert k.tolist() != b["k"]
    sup = r["support"].copy()
    sup[np.argmax(sup)] -= 1
    assert sup.tolist() != b["support"]


@pytest.mark.parametrize("name", sorted(n for n in EXPECTED if "rows" in EXPECTED[n]))
@pytest.mark.parametrize("mode", ["host", "ondemand"])
def test_golden(name, mode):
    exp = EXPECTED[name]
    r = run_oracle(os.path.join(GOLDEN, name), mode)
    rows = [[int(a), int(b), int(c), int(k)] for (a, b, c), k in zip(r["tris"].tolist(), r["k"].tolist())]
    assert r["T"] == exp["T"] and r["Q"] == exp["Q"] and r["K_max"] == exp["K_max"]
    assert rows == exp["rows"]


@pytest.mark.parametrize("parts", [1, 3, 8])
def test_parts_independent(tmp_path, parts):
    p = gen("rmat", 10, 8, 0.57, 0.19, 0.19, 21, out=tmp_path / "r.edges")
    a = run_oracle(p, "host", 1)
    b = run_oracle(p, "host", parts)
    assert np.array_equal(a["tris"], b["tris"]) and np.array_equal(a["k"], b["k"])


REF = os.path.join(os.path.dirname(oracle34.HERE), "oracle", "_ref", "ref34")


@pytest.mark.skipif(not os.path.exists(REF), reason="reference host path not built (oracle/_ref)")
@pytest.mark.parametrize("args", [("rmat", 11, 16, 0.57, 0.19, 0.19, 4), ("ba", 3000, 8, 5),
                                  ("er-deg", 4000, 24, 6), ("rmat", 10, 16, 0.45, 0.15, 0.15, 7)])
def test_against_reference_host_path(tmp_path, args):
    p = gen(*args, out=tmp_path / "g.edges")
    kf = tmp_path / "ref.k"
    subprocess.run([REF, p, str(kf), "4"], check=True, stdout=subprocess.DEVNULL)
    ref_rows = kf.read_text()
    for mode in ("host", "ondemand"):
        r = run_oracle(p, mode, 4)
        rows = "".join(f"({a},{b},{c})\t{k}\n" for (a, b, c), k in zip(r["tris"].tolist(), r["k"].tolist()))
        assert rows == ref_rows, mode
