"""Writes the golden fixtures under tests/golden/.

Inputs are small datasets covering the loader's quirks and the decomposition's
edge cases; expected outputs come from the oracle (oracle/), and every expected
output is re-derived here two more ways before it is written:
  * tests/brute.py (independent textbook restatement) for every fixture;
  * the reference's own host path (oracle/_ref/ref34, compiled from the
    reference headers) when that binary is present.
Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import struct
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import brute  # noqa: E402
import oracle34  # noqa: E402

GEN = os.path.join(ROOT, "gpu-nucleus34_amd", "_build", "nuc34_gen")


def write(name, text):
    with open(os.path.join(HERE, name), "w", newline="") as f:
        f.write(text)


def clique_lines(vs):
    return "".join(f"{a} {b}\n" for i, a in enumerate(vs) for b in vs[i + 1:])


def make_inputs():
    write("k5.edges", "0 0\n" + clique_lines([1, 2, 3, 4, 5]))
    write("two_cliques.edges", "0 0\n" + clique_lines(list(range(1, 7))) + clique_lines(list(range(10, 18))))
    hub = 0
    rim = list(range(1, 7))
    write("wheel.edges", "0 0\n" + "".join(f"{hub} {r}\n" for r in rim)
          + "".join(f"{rim[i]} {rim[(i + 1) % 6]}\n" for i in range(6)))
    write("star.edges", "0 0\n" + "".join(f"0 {i}\n" for i in range(1, 9)))
    # the first non-comment line of a bare list is dropped, even when it is an edge
    write("first_line.edges", "1 2\n" + clique_lines([1, 2, 3, 4]) .replace("1 2\n", "", 1) + "2 3\n1 3\n")
    write("quirks.txt",
          "# comment line\n% another comment\n0 0\n"
          "1 2 7.5\r\n2 1\r\n3 3\n1 3\n3 2 0.25\n2 3\n4 1\n4 2\n4 3\n5 5\n"
          "6 4\n6 1\n6 2\n6 3\n")
    write("onebased.mtx", "%%MatrixMarket matrix coordinate pattern symmetric\n% comment\n6 6 11\n"
          "1 2\n1 3\n1 4\n2 3\n2 4\n3 4\n4 5\n5 6\n4 6\n3 5\n3 6\n")
    write("twotok.mtx", "5 10\n" + clique_lines([0, 1, 2, 3, 4]))
    write("badheader.mtx", "% only a bad header follows\n7\n1 2\n")
    write("empty.edges", "0 0\n")
    write("unknown.xyz", "1 2\n")
    for name, args in [
        ("er40.edges", ["er-p", 40, 0.3, 7]),
        ("rmat8.edges", ["rmat", 8, 8, 0.57, 0.19, 0.19, 3]),
        ("ba60.edges", ["ba", 60, 5, 11]),
        ("rmat7k8.edges", ["rmat", 7, 8, 0.57, 0.19, 0.19, 5]),  # plus 2 planted K8
    ]:
        cmd = [GEN, *map(str, args), os.path.join(HERE, name)]
        if name == "rmat7k8.edges":
            cmd += ["2", "8"]
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    # .grh: an oriented graph stored as size_t 4 | nV | nE | D[nV] | E[nE]
    g = oracle34.OracleGraph.load(os.path.join(HERE, "er40.edges"), orient=True)
    O, E = g.csr()
    D = [int(O[i + 1] - O[i]) for i in range(len(O) - 1)]
    with open(os.path.join(HERE, "er40.grh"), "wb") as f:
        f.write(struct.pack("<QQQ", 4, len(D), len(E)))
        f.write(struct.pack(f"<{len(D)}I", *D))
        f.write(struct.pack(f"<{len(E)}I", *[int(x) for x in E]))


def parse_pairs(path):
    """Bare-list / mtx pair reading for the brute force (header / first line
    skipped, comments skipped, extra columns ignored)."""
    pairs = []
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines) and lines[i][:1] in ("#", "%"):
        i += 1
    for line in lines[i + 1:]:
        tok = line.split()
        if len(tok) >= 2:
            pairs.append((int(tok[0]), int(tok[1])))
    return pairs


def ref_rows(path):
    ref = os.path.join(ROOT, "oracle", "_ref", "ref34")
    if not os.path.exists(ref):
        return None
    with tempfile.TemporaryDirectory() as d:
        kf = os.path.join(d, "k")
        r = subprocess.run([ref, path, kf, "2"], stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        if r.returncode != 0:
            # e.g. the reference's host path crashes on a graph without edges
            return "crashed"
        rows = []
        with open(kf) as f:
            for line in f:
                t, k = line.rstrip("\n").split("\t")
                a, b, c = map(int, t.strip("()").split(","))
                rows.append([a, b, c, int(k)])
        return rows


def main():
    make_inputs()
    expected = {}
    for name in sorted(os.listdir(HERE)):
        path = os.path.join(HERE, name)
        ext = os.path.splitext(name)[1]
        if ext not in (".edges", ".txt", ".mtx", ".grh", ".xyz"):
            continue
        try:
            g = oracle34.OracleGraph.load(path, orient=False)
        except RuntimeError as e:
            expected[name] = {"error": str(e)}
            continue
        Ou, Eu = g.csr()
        go = oracle34.OracleGraph.load(path, orient=True)
        res = go.run("host", 2)
        od = go.run("ondemand", 2)
        rows = [[int(a), int(b), int(c), int(k)] for (a, b, c), k in zip(res["tris"].tolist(), res["k"].tolist())]
        assert od["T"] == res["T"] and od["Q"] == res["Q"] and (od["k"] == res["k"]).all(), name
        if ext != ".grh":
            n, nbrs = brute.undirected_from_pairs(parse_pairs(path))
            b = brute.decompose(n, nbrs)
            assert b["T"] == res["T"] and b["Q"] == res["Q"], name
            assert [list(t) + [k] for t, k in zip(b["tris"], b["k"])] == rows, name
            assert list(res["support"]) == b["support"], name
        rr = ref_rows(path)
        if isinstance(rr, list):
            assert rr == rows, f"reference host path disagrees on {name}"
        expected[name] = {
            "n": int(len(Ou) - 1), "m_undirected": int(len(Eu)),
            "und_O": [int(x) for x in Ou], "und_E": [int(x) for x in Eu],
            "T": int(res["T"]), "Q": int(res["Q"]), "K_max": int(res["K_max"]),
            "rows": rows, "reference_host_path": "agrees" if isinstance(rr, list) else (rr or "not built"),
        }
        if ext == ".grh":
            expected[name].pop("und_O")
            expected[name].pop("und_E")
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump(expected, f, separators=(",", ":"))
    print(f"wrote {len(expected)} fixtures")


if __name__ == "__main__":
    main()
