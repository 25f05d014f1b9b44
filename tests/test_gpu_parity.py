"""Parity of the HIP path (through the C ABI) with the oracle and the golden
vectors. Integer results must be bit-exact: triangle identities and order,
support, T, Q, K_max and every k."""
import hashlib
import itertools
import json
import os
import subprocess

import numpy as np
import pytest

import nuc34
import oracle34
from conftest import BUILD, GOLDEN, gen

pytestmark = pytest.mark.gpu
EXPECTED = json.load(open(os.path.join(GOLDEN, "expected.json")))
HASHES = json.load(open(os.path.join(GOLDEN, "hashes.json")))


def kfile_sha(tris, k):
    h = hashlib.sha256()
    chunk = 1 << 20
    for i in range(0, len(k), chunk):
        t = tris[i:i + chunk].tolist()
        kk = k[i:i + chunk].tolist()
        h.update("".join(f"({a},{b},{c})\t{x}\n" for (a, b, c), x in zip(t, kk)).encode())
    return h.hexdigest()


def gpu_run(g, variant, parts=1, budget=0):
    with nuc34.Context(0, budget) as ctx:
        ctx.load_graph(g)
        st = ctx.decompose(variant, parts)
        return ctx.copy_triangles(), ctx.copy_support(), ctx.copy_k(), st


@pytest.mark.parametrize("name", sorted(n for n in EXPECTED if "rows" in EXPECTED[n]))
@pytest.mark.parametrize("variant,parts", [("direct", 1), ("ondemand", 1), ("direct", 3), ("auto", 2)])
def test_golden(name, variant, parts):
    print(f"[golden] {name} {variant} parts={parts}", flush=True)
    exp = EXPECTED[name]
    g = nuc34.load_dataset(os.path.join(GOLDEN, name))
    tris, sup, k, st = gpu_run(g, variant, parts)
    rows = [[int(a), int(b), int(c), int(x)] for (a, b, c), x in zip(tris.tolist(), k.tolist())]
    assert st["triangles"] == exp["T"] and st["four_cliques"] == exp["Q"]
    assert rows == exp["rows"]
    if exp["Q"]:
        assert st["k_max"] == exp["K_max"]
    assert st["unpeeled"] == 0


CASES = [
    ("er-p", 300, 0.2, 1), ("er-deg", 3000, 40, 2), ("ba", 4000, 12, 3), ("ba", 2000, 3, 4),
    ("rmat", 11, 16, 0.57, 0.19, 0.19, 5), ("rmat", 12, 16, 0.57, 0.19, 0.19, 6),
    ("rmat", 12, 32, 0.45, 0.15, 0.15, 7), ("rmat", 10, 16, 0.57, 0.19, 0.19, 8, 6, 25),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c)))
def test_generated_vs_oracle(tmp_path, case):
    print(f"[generated] {case}", flush=True)
    p = gen(*case[:7], out=tmp_path / "g.edges") if case[0] != "rmat" or len(case) == 7 else None
    if p is None:
        subprocess.run([os.path.join(BUILD, "nuc34_gen"), *map(str, case[:7]), str(tmp_path / "g.edges"),
                        *map(str, case[7:])], check=True, stdout=subprocess.DEVNULL)
        p = str(tmp_path / "g.edges")
    ref = oracle34.OracleGraph.load(p, orient=True).run("ondemand", 8)
    g = nuc34.load_dataset(p)
    for variant in ("direct", "ondemand"):
        tris, sup, k, st = gpu_run(g, variant, parts=1 if variant == "direct" else 4)
        assert st["triangles"] == ref["T"] and st["four_cliques"] == ref["Q"], variant
        assert np.array_equal(tris, ref["tris"]), variant
        assert np.array_equal(sup, ref["support"]), variant
        assert np.array_equal(k, ref["k"]), variant
        assert st["k_max"] == ref["K_max"] and st["rounds"] == ref["rounds"], variant


def test_device_orientation_matches_reference_order(tmp_path):
    p = gen("rmat", 12, 16, 0.57, 0.19, 0.19, 9, out=tmp_path / "g.edges")
    g = nuc34.load_dataset(p)
    with nuc34.Context() as ctx:
        ctx.load_graph(g)
        O, E = ctx.oriented_csr()
    Oo, Eo = oracle34.OracleGraph.load(p, orient=True).csr()
    assert np.array_equal(O, Oo) and np.array_equal(E, Eo)


@pytest.mark.parametrize("n", [4, 6, 11, 40, 400])
def test_clique_closed_form(n):
    """K_n: C(n,3) triangles, C(n,4) K4s, k = n-3 everywhere, one round. K_400
    has 1.05 G K4s (more than 2^30), every one with all four members dying in
    the same round."""
    pairs = np.array(list(itertools.combinations(range(n), 2)), dtype=np.uint32)
    g = nuc34.graph_from_pairs(pairs[:, 0], pairs[:, 1])
    for variant in ("direct", "ondemand"):
        tris, sup, k, st = gpu_run(g, variant)
        assert st["triangles"] == n * (n - 1) * (n - 2) // 6
        assert st["four_cliques"] == n * (n - 1) * (n - 2) * (n - 3) // 24
        assert (k == n - 3).all() and st["k_max"] == n - 3 and st["rounds"] == 1


def test_no_triangles_and_empty():
    for pairs in ([(0, i) for i in range(1, 50)], [(0, 1)]):
        a = np.array(pairs, dtype=np.uint32)
        g = nuc34.graph_from_pairs(a[:, 0], a[:, 1])
        with nuc34.Context() as ctx:
            ctx.load_graph(g)
            assert ctx.triangles() == 0
            st = ctx.decompose("auto")
            assert st["triangles"] == 0 and st["four_cliques"] == 0 and len(ctx.copy_k()) == 0


def test_budget_selects_ondemand_and_reports_oom(tmp_path):
    p = gen("rmat", 12, 16, 0.57, 0.19, 0.19, 10, out=tmp_path / "g.edges")
    g = nuc34.load_dataset(p)
    with nuc34.Context(0, 64 << 20) as ctx:  # room for O(T) state, not the K4 list
        ctx.load_graph(g)
        st = ctx.decompose("auto")
        assert st["variant"] == nuc34.ONDEMAND
        k_small = ctx.copy_k()
    _, _, k_big, st2 = gpu_run(g, "direct")
    assert st2["variant"] == nuc34.DIRECT and np.array_equal(k_small, k_big)
    with nuc34.Context(0, 1 << 16) as ctx:
        with pytest.raises(nuc34.Nuc34Error) as ei:
            ctx.load_graph(g)
            ctx.decompose("direct")
        assert ei.value.code == nuc34.E_OOM and "budget" in str(ei.value)


@pytest.mark.parametrize("preset", ["er1000", "rmat16"])
def test_full_size_hash(tmp_path, preset):
    """Full configurations: the k file must hash to the value the reference's
    own host path produced; supports must sum to 4Q; k <= support."""
    p = gen("preset", preset, out=tmp_path / f"{preset}.edges")
    exp = HASHES[preset]
    g = nuc34.load_dataset(p)
    ks = {}
    for variant, parts in (("direct", 1), ("ondemand", 1), ("direct", 4)):
        print(f"[full] {preset} {variant} parts={parts}", flush=True)
        tris, sup, k, st = gpu_run(g, variant, parts)
        assert st["triangles"] == exp["T"] and st["four_cliques"] == exp["Q"] and st["k_max"] == exp["K_max"]
        assert int(sup.astype(np.int64).sum()) == 4 * exp["Q"]
        assert (k <= sup).all() and (k >= 0).all()
        assert kfile_sha(tris, k) == exp["kfile_sha256"], (variant, parts)
        ks[(variant, parts)] = k
    first = next(iter(ks.values()))
    assert all(np.array_equal(first, v) for v in ks.values())


def test_full_size_rmat20k30(tmp_path):
    """R-MAT scale 20 with 64 planted K30 (423 M triangles, 17.0 G four-cliques),
    on-demand: the raw k and triangle arrays hash to the oracle's, and T, Q,
    K_max and the number of peel rounds agree."""
    exp = HASHES["rmat20k30"]
    p = gen("preset", "rmat20k30", out=tmp_path / "rmat20k30.edges")
    g = nuc34.load_dataset(p)
    print("[full] rmat20k30 ondemand", flush=True)
    with nuc34.Context() as ctx:
        ctx.load_graph(g)
        del g
        st = ctx.decompose("ondemand")
        assert (st["triangles"], st["four_cliques"], st["k_max"]) == (exp["T"], exp["Q"], exp["K_max"])
        assert st["rounds"] == exp["rounds"] and st["unpeeled"] == 0
        k = ctx.copy_k()
        assert hashlib.sha256(k.tobytes()).hexdigest() == exp["k_raw_sha256"]
        del k
        tris = ctx.copy_triangles()
        assert hashlib.sha256(tris.tobytes()).hexdigest() == exp["tri_raw_sha256"]


def test_full_size_rmat18(tmp_path):
    """R-MAT scale 18 (the serving mix's large job: 83 M triangles, 2.3 G
    four-cliques), both variants: the raw k and triangle arrays hash to the
    oracle's, and T, Q, K_max and the peel rounds agree."""
    exp = HASHES["rmat18"]
    g = nuc34.load_dataset(gen("preset", "rmat18", out=tmp_path / "rmat18.edges"))
    for variant in ("ondemand", "direct"):
        print(f"[full] rmat18 {variant}", flush=True)
        with nuc34.Context() as ctx:
            ctx.load_graph(g)
            st = ctx.decompose(variant)
            assert (st["triangles"], st["four_cliques"], st["k_max"]) == (exp["T"], exp["Q"], exp["K_max"])
            assert st["rounds"] == exp["rounds"] and st["unpeeled"] == 0
            assert hashlib.sha256(ctx.copy_k().tobytes()).hexdigest() == exp["k_raw_sha256"], variant
            assert hashlib.sha256(ctx.copy_triangles().tobytes()).hexdigest() == exp["tri_raw_sha256"], variant


def test_cli_matches_golden(tmp_path):
    for exe, name in (("nucleus34", "rmat8.edges"), ("nucleus34_ondemand", "rmat7k8.edges"),
                      ("nucleus34", "er40.grh"), ("nucleus34", "star.edges"), ("nucleus34_ondemand", "wheel.edges")):
        exp = EXPECTED[name]
        out = tmp_path / f"{name}.k"
        args = [os.path.join(BUILD, exe), os.path.join(GOLDEN, name)]
        args += (["2", str(out)] if exe == "nucleus34" else [str(out)])
        r = subprocess.run(args, capture_output=True, text=True, check=True)
        rec = json.loads(r.stdout)["dataset"]
        assert "error" not in rec, rec
        assert rec["triangle_count"] == exp["T"]
        if exp["T"] == 0:
            assert "four_clique_count" not in rec
            continue
        assert rec["four_clique_count"] == exp["Q"]
        if exp["Q"] == 0:
            assert "K_max" not in rec
            continue
        assert rec["K_max"] == exp["K_max"] and rec["K_count"] == exp["T"]
        assert rec["K"] == [r_[3] for r_ in exp["rows"][:5]]
        rows = [[*map(int, line.split("\t")[0].strip("()").split(",")), int(line.split("\t")[1])]
                for line in out.read_text().splitlines()]
        assert rows == exp["rows"]


def test_graph_rounds_match_eager_rounds(tmp_path, monkeypatch):
    """Peel rounds run as a captured graph replayed many times, with host syncs
    between replays; every replayed run must equal the eagerly launched one and
    the oracle."""
    p = gen("rmat", 12, 16, 0.57, 0.19, 0.19, 12, out=tmp_path / "g.edges")
    # the on-demand restatement's rounds are the level-synchronous peel's (the
    # host path's bucket peel counts its batches differently); both give the same k
    ref = oracle34.OracleGraph.load(p, orient=True).run("ondemand", 8)
    assert np.array_equal(ref["k"], oracle34.OracleGraph.load(p, orient=True).run("host", 8)["k"])
    g = nuc34.load_dataset(p)
    for variant in ("direct", "ondemand"):
        monkeypatch.setenv("NUC34_NO_GRAPH", "1")
        _, _, k_eager, st_eager = gpu_run(g, variant)
        monkeypatch.delenv("NUC34_NO_GRAPH")
        assert np.array_equal(k_eager, ref["k"]) and st_eager["rounds"] == ref["rounds"], variant
        for _ in range(3):
            _, _, k_graph, st_graph = gpu_run(g, variant)
            assert np.array_equal(k_graph, k_eager), variant
            assert st_graph["rounds"] == st_eager["rounds"] and st_graph["k_max"] == st_eager["k_max"]
        assert np.array_equal(k_eager, ref["k"]), variant


@pytest.mark.parametrize("name", ["rmat8.edges", "rmat7k8.edges", "two_cliques.edges"])
def test_k4_list(name):
    """The direct variant's K4 list: every K4 exactly once, as
    {T(pqs), T(pqr), T(prs), T(qrs)} for p<q<r<s in rank order, grouped by the
    triangle on the three lowest vertices (ascending)."""
    import brute
    from test_golden_helpers import pairs_of
    path = os.path.join(GOLDEN, name)
    n, nbrs = brute.undirected_from_pairs(pairs_of(path))
    b = brute.decompose(n, nbrs)
    g = nuc34.load_dataset(path)
    with nuc34.Context() as ctx:
        ctx.load_graph(g)
        st = ctx.decompose("direct")
        quds = ctx.copy_fourcliques()
        tris = ctx.copy_triangles()
        O, E = ctx.oriented_csr()
    assert len(quds) == b["Q"] == st["four_cliques"]
    assert {frozenset(map(int, q)) for q in quds} == set(b["k4s"])
    deg = np.array([len(x) for x in nbrs])
    src = np.repeat(np.arange(len(O) - 1), np.diff(O))
    def verts(t):
        a, bb, c = tris[t]
        return {int(src[bb]), int(E[bb]), int(E[a])}
    rank = lambda v: (deg[v], v)
    for q in quds:
        vs = sorted(set().union(*(verts(int(t)) for t in q)), key=rank)
        p_, q_, r_, s_ = vs
        assert [verts(int(t)) for t in q] == [{p_, q_, s_}, {p_, q_, r_}, {p_, r_, s_}, {q_, r_, s_}]
    assert all(quds[i][1] <= quds[i + 1][1] for i in range(len(quds) - 1))


@pytest.mark.parametrize("variant", ["direct", "ondemand"])
def test_batch_matches_single_graphs(variant):
    """A batch is decomposed as one disjoint union: every graph's slice must
    equal its own decomposition, and the smallest ones the oracle's."""
    idx = list(range(9)) + [30, 31, 32]
    gs = [nuc34.gen_batch_graph(3, i) for i in idx]
    with nuc34.Context() as ctx:
        ctx.load_batch(gs)
        st = ctx.decompose(variant)
        tris = ctx.copy_triangles()
        k = ctx.copy_k()
        sup = ctx.copy_support()
        infos = [ctx.batch_graph(g) for g in range(len(gs))]
    assert st["unpeeled"] == 0
    assert sum(i["tri_count"] for i in infos) == len(k)
    order = sorted(range(len(gs)), key=lambda j: gs[j].m)
    for j, g in enumerate(gs):
        info = infos[j]
        sl = slice(info["tri_first"], info["tri_first"] + info["tri_count"])
        t_single, s_single, k_single, st_single = gpu_run(g, variant)
        assert np.array_equal(tris[sl] - np.uint32(info["edge_first"]), t_single), j
        assert np.array_equal(k[sl], k_single), j
        assert np.array_equal(sup[sl], s_single), j
        if j in order[:3]:
            ref = oracle34.OracleGraph.from_csr(g.O, g.E, orient=True).run("host", 4)
            assert np.array_equal(k_single, ref["k"]) and np.array_equal(t_single, ref["tris"]), j


def test_full_batch_512_graphs_vs_oracle():
    """The whole many-graphs workload (512 graphs of the seeded mix, seed 3) as
    one disjoint union, in both variants: every graph's slice of triangles,
    support and k equals the oracle's decomposition of that graph alone."""
    gs = [nuc34.gen_batch_graph(3, i) for i in range(512)]
    refs = [oracle34.OracleGraph.from_csr(g.O, g.E, orient=True).run("ondemand", 16) for g in gs]
    for variant, batch_peel in itertools.product(("ondemand", "direct"), (2, 1)):
        print(f"[batch512] {variant} batch_peel={batch_peel}", flush=True)
        with nuc34.Context() as ctx:
            ctx.set_tuning(batch_peel=batch_peel)
            ctx.load_batch(gs)
            st = ctx.decompose(variant)
            tris, k, sup = ctx.copy_triangles(), ctx.copy_k(), ctx.copy_support()
            ef, tf = ctx.batch_slices()
            rounds = ctx.batch_rounds() if batch_peel == 2 else None
        assert st["unpeeled"] == 0 and st["triangles"] == sum(r["T"] for r in refs)
        assert st["four_cliques"] == sum(r["Q"] for r in refs)
        assert st["k_max"] == max(r["K_max"] for r in refs)
        if batch_peel == 2:  # one workgroup per graph: every graph's own rounds
            assert [int(x) for x in rounds] == [r["rounds"] for r in refs], variant
            assert st["rounds"] == max(r["rounds"] for r in refs)
        else:
            assert st["rounds"] >= max(r["rounds"] for r in refs)
        for j, r in enumerate(refs):
            sl = slice(int(tf[j]), int(tf[j + 1]))
            assert np.array_equal(tris[sl] - np.uint32(ef[j]), r["tris"]), (variant, j)
            assert np.array_equal(sup[sl], r["support"]), (variant, j)
            assert np.array_equal(k[sl], r["k"]), (variant, j)


def _pairs(edges):
    a = np.array([x for x, _ in edges], dtype=np.uint32)
    b = np.array([y for _, y in edges], dtype=np.uint32)
    return nuc34.graph_from_pairs(a, b)


@pytest.mark.parametrize("variant", ["direct", "ondemand"])
def test_batch_peel_one_workgroup_per_graph(variant):
    """A batch peeled one workgroup per graph (the whole peel one launch) gives
    the k of the union's shared rounds, and every graph's rounds equal the
    oracle's for that graph alone. Graphs without triangles (a star, no edges),
    without K4s (a wheel) and single cliques (one round, k = n - 3) sit between
    mixed-batch graphs, so graph slices start and end inside state words."""
    star = _pairs([(0, i) for i in range(1, 9)])
    wheel = _pairs([(0, i) for i in range(1, 9)] + [(i, i % 8 + 1) for i in range(1, 9)])
    k7 = _pairs(list(itertools.combinations(range(7), 2)))
    k12 = _pairs(list(itertools.combinations(range(12), 2)))
    empty = _pairs([(0, 0)])
    mix = [nuc34.gen_batch_graph(3, i) for i in range(20)]
    gs = [mix[0], star, mix[1], wheel, k7, mix[2], empty, k12] + mix[3:] + [wheel, k7]
    refs = [oracle34.OracleGraph.from_csr(g.O, g.E, orient=True).run("ondemand", 4) for g in gs]
    out = {}
    for batch_peel in (2, 1):
        with nuc34.Context() as ctx:
            ctx.set_tuning(batch_peel=batch_peel)
            ctx.load_batch(gs)
            st = ctx.decompose(variant)
            out[batch_peel] = (ctx.copy_k(), st, ctx.batch_slices()[1],
                               ctx.batch_rounds() if batch_peel == 2 else None)
            if batch_peel == 1:
                with pytest.raises(nuc34.Nuc34Error):
                    ctx.batch_rounds()
    k2, st2, tf, rounds = out[2]
    k1, st1, _, _ = out[1]
    assert np.array_equal(k2, k1) and st2["unpeeled"] == 0 and st2["k_max"] == st1["k_max"]
    assert [int(x) for x in rounds] == [r["rounds"] for r in refs]
    assert st2["rounds"] == max(r["rounds"] for r in refs)
    for j, r in enumerate(refs):
        assert np.array_equal(k2[int(tf[j]):int(tf[j + 1])], r["k"]), j
    assert rounds[7] == 1 and (k2[int(tf[7]):int(tf[8])] == 9).all()  # gs[7] is K_12


def test_auto_variant_for_batches():
    """AUTO runs a batch of small graphs direct (their lists fit) and a batch
    holding a graph above graph_peel_max_tris on demand; one graph on demand."""
    gs = [nuc34.gen_batch_graph(3, i) for i in range(4)]
    with nuc34.Context() as ctx:
        ctx.load_batch(gs)
        assert ctx.decompose("auto")["variant"] == nuc34.DIRECT
        k_small = ctx.copy_k()
        ctx.triangles()
        most = max(ctx.batch_graph(g)["tri_count"] for g in range(len(gs)))
        ctx.set_tuning(graph_peel_max_tris=max(most - 1, 1))
        ctx.reset()
        assert ctx.decompose("auto")["variant"] == nuc34.ONDEMAND
        assert np.array_equal(ctx.copy_k(), k_small)
        ctx.load_graph(gs[0])
        assert ctx.decompose("auto")["variant"] == nuc34.ONDEMAND


def test_batch_peel_falls_back_to_shared_rounds_for_large_graphs():
    """A graph above graph_peel_max_tris keeps the batch on shared rounds (no
    per-graph rounds), with the same k; raising the limit peels it by graph."""
    gs = [nuc34.gen_batch_graph(3, i) for i in range(6)]
    with nuc34.Context() as ctx:
        ctx.load_batch(gs)
        ctx.triangles()
        tc = max(ctx.batch_graph(g)["tri_count"] for g in range(len(gs)))
        ctx.set_tuning(graph_peel_max_tris=max(tc - 1, 1))
        ctx.decompose("direct")
        k_shared = ctx.copy_k()
        with pytest.raises(nuc34.Nuc34Error):
            ctx.batch_rounds()
        ctx.set_tuning(graph_peel_max_tris=tc)
        ctx.reset()
        ctx.decompose("direct")
        assert np.array_equal(ctx.copy_k(), k_shared) and len(ctx.batch_rounds()) == len(gs)


def test_fourclique_copy_fits_beside_the_direct_peel():
    """Copying the first K4s after a direct decomposition needs O(T) more
    device memory, not O(Q): it succeeds with a budget just above the
    decomposition's peak (a dense graph, so 8 bytes per K4 would not fit), and
    the K4s equal those of an unconstrained context."""
    n = 70
    pairs = np.array([(i, j) for i in range(n) for j in range(i + 1, n) if (i * 7 + j * 3) % 11], dtype=np.uint32)
    g = nuc34.graph_from_pairs(pairs[:, 0], pairs[:, 1])
    with nuc34.Context(0) as ctx:
        ctx.load_graph(g)
        st = ctx.decompose("direct")
        want = ctx.copy_fourcliques(0, 5)
        mid = ctx.copy_fourcliques(st["four_cliques"] // 2, 7)
    T, Q = st["triangles"], st["four_cliques"]
    assert 8 * Q > 64 * (T + 1) + (1 << 20)
    budget = st["peak_bytes"] + 24 * (T + 1) + (256 << 10)
    with nuc34.Context(0, budget) as ctx:
        ctx.load_graph(g)
        ctx.decompose("direct")
        assert np.array_equal(ctx.copy_fourcliques(0, 5), want)
        assert np.array_equal(ctx.copy_fourcliques(Q // 2, 7), mid)


def test_batch_graph_index_bounds():
    """nuc34_batch_graph rejects indices past the last graph, 2^32 - 1 included
    (the index is compared in 64 bits)."""
    import ctypes as C
    gs = [nuc34.gen_batch_graph(3, i) for i in range(2)]
    with nuc34.Context() as ctx:
        ctx.load_batch(gs)
        ctx.decompose("auto")
        vf = C.c_uint64()
        for g in (2, 0xFFFFFFFF):
            rc = nuc34._lib.nuc34_batch_graph(ctx._h, g, C.byref(vf), None, None, None)
            assert rc == nuc34.E_INVALID, g
        assert nuc34._lib.nuc34_batch_graph(ctx._h, 1, C.byref(vf), None, None, None) == nuc34.OK
        assert vf.value == gs[0].n


def test_batch_concatenated_and_pointer_loads_agree():
    """nuc34_load_batch (graphs back to back) and nuc34_load_batch_ptrs (one
    pointer pair per graph, an edgeless graph among them) load the same union."""
    import ctypes as C
    gs = [(np.asarray(g.O, np.uint32), np.asarray(g.E, np.uint32)) for g in
          (nuc34.gen_batch_graph(3, i) for i in range(6))]
    gs.insert(2, (np.zeros(4, np.uint32), np.zeros(0, np.uint32)))  # three isolated vertices
    res = []
    for mode in ("concatenated", "pointers"):
        with nuc34.Context() as ctx:
            if mode == "pointers":
                ctx.load_batch(gs)
            else:
                n = np.array([len(O) - 1 for O, _ in gs], np.uint32)
                m = np.array([len(E) for _, E in gs], np.uint64)
                O = np.concatenate([O for O, _ in gs])
                E = np.concatenate([E for _, E in gs])
                rc = nuc34._lib.nuc34_load_batch(ctx._h, len(gs), nuc34._u32(n),
                                                 m.ctypes.data_as(C.POINTER(C.c_uint64)), nuc34._u32(O),
                                                 nuc34._u32(E), 0)
                assert rc == nuc34.OK
                ctx._batch = len(gs)
            st = ctx.decompose("auto")
            res.append((ctx.copy_triangles(), ctx.copy_k(), ctx.batch_slices(), st["four_cliques"]))
    (t0, k0, s0, q0), (t1, k1, s1, q1) = res
    assert q0 == q1 and np.array_equal(t0, t1) and np.array_equal(k0, k1)
    assert all(np.array_equal(a, b) for a, b in zip(s0, s1))


@pytest.mark.parametrize("case", ["offsets", "range", "self_loop", "unsorted", "asymmetric"])
def test_malformed_csr_rejected(case):
    """Malformed undirected input is rejected by the device check (which reads
    only inside the arrays) with E_INVALID, alone and inside a batch, and the
    context stays usable."""
    good = (np.array([0, 2, 4, 6], np.uint32), np.array([1, 2, 0, 2, 0, 1], np.uint32))  # a triangle
    bad = {
        "offsets": (np.array([0, 3, 2, 6], np.uint32), np.array([1, 2, 0, 2, 0, 1], np.uint32)),
        "range": (np.array([0, 2, 4, 6], np.uint32), np.array([1, 7, 0, 2, 0, 1], np.uint32)),
        "self_loop": (np.array([0, 2, 4, 6], np.uint32), np.array([0, 2, 0, 2, 0, 1], np.uint32)),
        "unsorted": (np.array([0, 2, 4, 6], np.uint32), np.array([2, 1, 0, 2, 0, 1], np.uint32)),
        "asymmetric": (np.array([0, 1, 3, 4], np.uint32), np.array([1, 0, 2, 0], np.uint32)),
    }[case]
    print(f"[malformed] {case}", flush=True)
    with nuc34.Context() as ctx:
        with pytest.raises(nuc34.Nuc34Error) as ei:
            ctx.load_csr(bad[0], bad[1], False)
        assert ei.value.code == nuc34.E_INVALID
        with pytest.raises(nuc34.Nuc34Error) as ei:
            ctx.load_batch([good, bad, good])
        assert ei.value.code == nuc34.E_INVALID
        if case != "asymmetric":  # found by the check itself, which names the graph
            assert "graph 1" in str(ei.value)
        ctx.load_batch([good, good])
        st = ctx.decompose("auto")
        assert st["triangles"] == 2 and st["four_cliques"] == 0


def _hub_core_graph(seed=7):
    """900 hubs with 870 private leaves each and about 40 random hub neighbours;
    X sees all 900 hubs (out-degree 900 > 640: edge tiles), Y the first 400
    (a 1024-thread workgroup), Z the first 200 (a 256-thread workgroup), X-Y is
    an edge; hubs (about 20 out-neighbours) and leaves take a wave each."""
    rng = np.random.default_rng(seed)
    H, L = 900, 870
    X, Y, Z = H, H + 1, H + 2
    src, dst = [], []
    for h in range(H):
        leaves = np.arange(H + 3 + h * L, H + 3 + (h + 1) * L)
        src += [h] * L
        dst += leaves.tolist()
    a = rng.integers(0, H, 18000)
    b = rng.integers(0, H, 18000)
    src += a.tolist()
    dst += b.tolist()
    src += [X] * H + [Y] * 400 + [Z] * 200 + [X]
    dst += list(range(H)) + list(range(400)) + list(range(200)) + [Y]
    return nuc34.graph_from_pairs(np.array(src, dtype=np.uint32), np.array(dst, dtype=np.uint32))


@pytest.mark.parametrize("variant,parts", [("direct", 1), ("ondemand", 1), ("direct", 3), ("ondemand", 2)])
def test_degree_binned_intersection_matches_tiles(variant, parts):
    """The degree-binned intersection (tri_schedule 1: a wave, a 256-thread or
    a 1024-thread workgroup per vertex by out-degree, edge tiles above 640; 2:
    also tiles up to out-degree 8) gives the edge tiles' triangle list,
    support, k and rounds, on R-MAT s16
    (owners up to out-degree 248) and on a hub-core graph with owners in every
    class; the hub-core results also equal the oracle's."""
    hub = _hub_core_graph()
    ref = oracle34.OracleGraph.from_csr(hub.O, hub.E, orient=True).run("ondemand", 4)
    for name, g in (("hub-core", hub), ("rmat16", nuc34.gen_preset("rmat16"))):
        out = []
        for schedule in (0, 1, 2):
            print(f"[binned] {name} {variant} parts={parts} tri_schedule={schedule}", flush=True)
            with nuc34.Context() as ctx:
                ctx.set_tuning(tri_schedule=schedule)
                ctx.load_graph(g)
                st = ctx.decompose(variant, parts)
                out.append((ctx.copy_triangles(), ctx.copy_support(), ctx.copy_k(), st))
        t0, s0, k0, st0 = out[0]
        for t1, s1, k1, st1 in out[1:]:
            assert np.array_equal(t0, t1) and np.array_equal(s0, s1) and np.array_equal(k0, k1), name
            assert st0["rounds"] == st1["rounds"] and st0["four_cliques"] == st1["four_cliques"], name
        if name == "hub-core":
            assert np.array_equal(t1, ref["tris"]) and np.array_equal(k1, ref["k"]) and st1["rounds"] == ref["rounds"]


def test_autotune_and_launch_shapes_keep_results(tmp_path):
    """Launch shapes change speed, never results: the autotuned shape and odd
    explicit shapes give the same k and round count in both variants."""
    p = gen("rmat", 12, 16, 0.57, 0.19, 0.19, 11, out=tmp_path / "g.edges")
    g = nuc34.load_dataset(p)
    with nuc34.Context() as ctx:
        ctx.load_graph(g)
        st0 = ctx.decompose("ondemand")
        k0 = ctx.copy_k()
        t, ms = ctx.autotune("ondemand")
        print(f"[autotune] {t} {ms:.3f} ms", flush=True)
        assert ms > 0 and 1 <= t["peel_tile_cap"] <= 64 and t["peel_rounds_per_launch"] % 2 == 0
        for variant, shape in (("ondemand", {}), ("ondemand", dict(peel_debit_grid=96, peel_tile_cap=3,
                                                                    peel_rounds_per_launch=6)),
                               ("direct", dict(peel_debit_grid=300, k4_list_grid_cap=17)),
                               ("direct", dict(tri_tile_edges=16)), ("ondemand", dict(tri_tile_edges=32)),
                               ("ondemand", dict(tri_schedule=1)), ("direct", dict(tri_schedule=0)),
                               # two debit workgroups: their push regions (1024 slots) overflow into
                               # the shared list in the rounds with more pushes
                               ("ondemand", dict(peel_debit_grid=2, peel_tile_cap=64)),
                               ("direct", dict(peel_debit_grid=2))):
            if shape:
                ctx.set_tuning(**shape)
            ctx.reset()
            st = ctx.decompose(variant)
            assert np.array_equal(ctx.copy_k(), k0) and st["rounds"] == st0["rounds"], (variant, shape)
        with pytest.raises(nuc34.Nuc34Error):
            ctx.set_tuning(peel_tile_cap=65)
        with pytest.raises(nuc34.Nuc34Error):
            ctx.set_tuning(tri_tile_edges=48)
        with pytest.raises(nuc34.Nuc34Error):
            ctx.set_tuning(tri_schedule=3)


def _clique_book(fans):
    """A triangle a, b, c and `fans` vertices adjacent to all three, in groups
    of 1, 2, 3, 4, 5, ... (cycling) that are cliques: each group of g with
    a, b, c is a K(g+3), so triangle abc has support `fans` and the other
    triangles k = g."""
    a, b, c = fans, fans + 1, fans + 2
    src, dst = [a, a, b], [b, c, c]
    f, g = 0, 1
    while f < fans:
        grp = list(range(f, min(f + g, fans)))
        for i, x in enumerate(grp):
            src += [x, x, x] + [x] * (len(grp) - i - 1)
            dst += [a, b, c] + grp[i + 1:]
        f += len(grp)
        g = g % 5 + 1
    return nuc34.graph_from_pairs(np.array(src, dtype=np.uint32), np.array(dst, dtype=np.uint32))


def test_hub_triangle_with_support_above_16_bits():
    """One triangle in 65536 K4s (its three vertices adjacent to every other
    vertex, so its down rows span 2049 words) beside many small cliques: both
    variants equal the oracle's triangles, k and rounds."""
    g = _clique_book(65536)
    ref = oracle34.OracleGraph.from_csr(g.O, g.E, orient=True).run("ondemand", 4)
    assert int(ref["support"].max()) == 65536
    for variant in ("ondemand", "direct"):
        print(f"[hub-triangle] {variant}", flush=True)
        with nuc34.Context() as ctx:
            ctx.load_graph(g)
            st = ctx.decompose(variant)
            assert np.array_equal(ctx.copy_triangles(), ref["tris"]) and np.array_equal(ctx.copy_k(), ref["k"])
            assert np.array_equal(ctx.copy_support(), ref["support"]), variant
            assert st["rounds"] == ref["rounds"] and st["k_max"] == ref["K_max"] and st["unpeeled"] == 0


def level_peel_sim(tris, O, E):
    """Synchronous level peel of the triangle/K4 hypergraph, brute force: each
    round removes every live triangle whose live support is at most the level,
    and each K4 that loses its first members debits its surviving members once.
    `tris` holds each triangle's three edge ids in the oriented CSR (O, E).
    Returns (k, rounds, debits, essential): `debits` counts every debit of a
    survivor, `essential` only those that bring a survivor down to the level
    (the rest may be skipped once it is pushed)."""
    n = len(O) - 1
    src = np.repeat(np.arange(n), np.diff(O.astype(np.int64)))
    adj = [set() for _ in range(n)]
    for u, v in zip(src.tolist(), E.tolist()):
        adj[u].add(v)
        adj[v].add(u)
    tid = {}
    for i, t in enumerate(tris.tolist()):
        tid[tuple(sorted({int(src[e]) for e in t} | {int(E[e]) for e in t}))] = i
    assert len(tid) == len(tris) and all(len(t) == 3 for t in tid)
    quads = []
    for a, b, c in tid:
        for x in adj[a] & adj[b] & adj[c]:
            if x > c:
                quads.append([tid[(a, b, c)], tid[(a, b, x)], tid[(a, c, x)], tid[(b, c, x)]])
    K = np.array(quads, dtype=np.int64).reshape(-1, 4)
    d = np.bincount(K.ravel(), minlength=len(tris)).astype(np.int64)
    alive = np.ones(len(tris), bool)
    dead = np.zeros(len(K), bool)
    k = np.zeros(len(tris), np.int64)
    level, rounds, debits, essential = 0, 0, 0, 0
    while alive.any():
        level = max(level, int(d[alive].min()))
        S = alive & (d <= level)
        while S.any():
            rounds += 1
            hit = ~dead & S[K].any(axis=1)
            dead |= hit
            members = K[hit].ravel()
            take = members[alive[members] & ~S[members]]
            r = np.bincount(take, minlength=len(tris))
            essential += int(np.minimum(r, np.maximum(d - level, 0)).sum())
            d -= r
            debits += len(take)
            k[S] = level
            alive &= ~S
            S = alive & (d <= level)
    return k, rounds, debits, essential


@pytest.mark.parametrize("name", ["rmat8.edges", "rmat7k8.edges", "ba60.edges", "two_cliques.edges"])
def test_peel_debit_count_and_rounds(name):
    """The peel's round count equals a brute-force synchronous level peel's,
    whose k and rounds agree with the oracle's; its debit count lies between
    the essential debits and all debits of that peel (a survivor pushed during
    a round may or may not see its later debits of the round, by timing)."""
    print(f"[debits] {name}", flush=True)
    g = nuc34.load_dataset(os.path.join(GOLDEN, name))
    og = oracle34.OracleGraph.from_csr(g.O, g.E, orient=True)
    ref = og.run("ondemand", 1)
    k, rounds, debits, essential = level_peel_sim(ref["tris"], *og.csr())
    assert np.array_equal(k, ref["k"]) and rounds == ref["rounds"]
    for variant in ("direct", "ondemand"):
        _, _, kg, st = gpu_run(g, variant)
        assert np.array_equal(kg, ref["k"]) and st["rounds"] == rounds, variant
        assert essential <= st["debits"] <= debits, (variant, st["debits"], essential, debits)


def test_device_rates():
    r = nuc34.device_rates(0)
    print(f"[rates] {r}", flush=True)
    assert all(np.isfinite(v) and v > 0 for v in r.values())
    assert r["copy_gbps"] > 100 and r["atomic_gops"] > 0.5


@pytest.mark.parametrize("variant", ["direct", "ondemand"])
def test_peel_grids_beyond_residency(tmp_path, variant):
    """Peel kernels with far more workgroups than the device holds at once
    (late workgroups start after early ones have ended) give the same k and
    rounds: no workgroup's control reads depend on another's writes."""
    p = gen("rmat", 12, 16, 0.57, 0.19, 0.19, 11, out=tmp_path / "g.edges")
    ref = oracle34.OracleGraph.load(p, orient=True).run("ondemand", 8)
    g = nuc34.load_dataset(p)
    with nuc34.Context() as ctx:
        ctx.load_graph(g)
        for shape in (dict(peel_finish_grid=1 << 16), dict(peel_finish_grid=1 << 15, peel_debit_grid=1 << 15),
                      dict(peel_finish_grid=1, peel_debit_grid=1)):
            print(f"[grids] {variant} {shape}", flush=True)
            ctx.set_tuning(**shape)
            ctx.reset()
            st = ctx.decompose(variant)
            assert np.array_equal(ctx.copy_k(), ref["k"]), shape
            assert st["rounds"] == ref["rounds"] and st["unpeeled"] == 0, shape


def test_concurrent_contexts_bit_exact(tmp_path):
    """Two contexts decomposing at the same time on their own streams (one
    large graph over and over, many small ones beside it) give each graph the
    result it gets alone."""
    import threading
    p = gen("rmat", 14, 16, 0.57, 0.19, 0.19, 5, out=tmp_path / "big.edges")
    big = nuc34.load_dataset(p)
    small = [nuc34.gen_batch_graph(4, i) for i in range(0, 40, 3)]
    alone = {}
    for i, g in enumerate(small):
        for variant in ("direct", "ondemand"):
            t, _, k, _ = gpu_run(g, variant)
            alone[(i, variant)] = (t, k)
    _, _, kbig, _ = gpu_run(big, "ondemand")
    stop, errors = threading.Event(), []

    def hog():
        with nuc34.Context() as ctx:
            ctx.load_graph(big)
            while not stop.is_set():
                ctx.reset()
                ctx.decompose("ondemand")
                if not np.array_equal(ctx.copy_k(), kbig):
                    errors.append("large graph")
                    return

    th = threading.Thread(target=hog)
    th.start()
    try:
        for rep in range(3):
            print(f"[concurrent] pass {rep}", flush=True)
            for i, g in enumerate(small):
                for variant in ("direct", "ondemand"):
                    t, _, k, _ = gpu_run(g, variant)
                    et, ek = alone[(i, variant)]
                    if not (np.array_equal(t, et) and np.array_equal(k, ek)):
                        errors.append((rep, i, variant))
    finally:
        stop.set()
        th.join()
    assert not errors, errors


def test_auto_fit_picks_by_footprint(tmp_path):
    """NUC34_AUTO_FIT runs direct when the K4 lists fit what is left of the
    budget and on demand when they do not; k is the same either way."""
    p = gen("rmat", 12, 16, 0.57, 0.19, 0.19, 10, out=tmp_path / "g.edges")
    g = nuc34.load_dataset(p)
    with nuc34.Context(0) as ctx:  # no budget: the device's free memory decides
        ctx.load_graph(g)
        st = ctx.decompose("auto_fit")
        assert st["variant"] == nuc34.DIRECT
        k_direct = ctx.copy_k()
    with nuc34.Context(0, 64 << 20) as ctx:  # O(T) state fits, the lists do not
        ctx.load_graph(g)
        st = ctx.decompose("auto_fit")
        assert st["variant"] == nuc34.ONDEMAND
        assert np.array_equal(ctx.copy_k(), k_direct)


def random_graph(rng, kind):
    """Small graphs of assorted shapes: sparse and dense ER, overlapping planted
    cliques, a hub with a dense neighbourhood, ids spread over a wide range."""
    if kind == "er":
        n = int(rng.integers(20, 300))
        p = float(rng.uniform(0.02, 0.4))
        iu = np.triu_indices(n, 1)
        keep = rng.random(len(iu[0])) < p
        a, b = iu[0][keep], iu[1][keep]
    elif kind == "cliques":
        n = int(rng.integers(30, 200))
        pairs = []
        for _ in range(int(rng.integers(2, 8))):
            vs = rng.choice(n, size=int(rng.integers(4, 14)), replace=False)
            pairs += [(x, y) for i, x in enumerate(vs) for y in vs[i + 1:]]
        pairs += [tuple(rng.choice(n, 2, replace=False)) for _ in range(n)]
        a, b = np.array(pairs).T
    elif kind == "hub":
        n = int(rng.integers(50, 250))
        hub_nbrs = rng.choice(np.arange(1, n), size=n // 2, replace=False)
        pairs = [(0, int(x)) for x in hub_nbrs]
        pairs += [tuple(rng.choice(hub_nbrs, 2, replace=False)) for _ in range(3 * n)]
        a, b = np.array(pairs).T
    else:  # spread ids: the same edges on ids scattered up to 50x the vertex count
        n = int(rng.integers(20, 120))
        iu = np.triu_indices(n, 1)
        keep = rng.random(len(iu[0])) < 0.3
        ids = np.sort(rng.choice(50 * n, size=n, replace=False))
        a, b = ids[iu[0][keep]], ids[iu[1][keep]]
    a, b = np.asarray(a, np.uint32), np.asarray(b, np.uint32)
    m = a != b
    return nuc34.graph_from_pairs(a[m], b[m])


@pytest.mark.parametrize("kind", ["er", "cliques", "hub", "spread"])
def test_random_small_graphs_vs_oracle(kind):
    """Many small random graphs per shape: triangles, support, k, K_max and
    rounds equal the oracle's in both variants and with parts > 1."""
    rng = np.random.default_rng({"er": 11, "cliques": 12, "hub": 13, "spread": 14}[kind])
    for it in range(12):
        g = random_graph(rng, kind)
        print(f"[random] {kind} #{it} n={g.n} m={g.m}", flush=True)
        ref = oracle34.OracleGraph.from_csr(g.O, g.E, orient=True).run("ondemand", 4)
        for variant, parts in (("direct", 1), ("ondemand", 1), ("direct", 3)):
            tris, sup, k, st = gpu_run(g, variant, parts)
            assert st["triangles"] == ref["T"] and st["four_cliques"] == ref["Q"], (kind, it, variant)
            assert np.array_equal(tris, ref["tris"]) and np.array_equal(sup, ref["support"]), (kind, it, variant)
            assert np.array_equal(k, ref["k"]), (kind, it, variant, parts)
            if ref["Q"]:
                assert st["k_max"] == ref["K_max"] and st["rounds"] == ref["rounds"], (kind, it, variant)


@pytest.mark.parametrize("exe", ["nucleus34", "nucleus34_ondemand"])
def test_cli_full_size_kfile(tmp_path, exe):
    """The drop-in CLIs at full size (R-MAT scale 16): the k file they write
    hashes to the reference host path's, and the JSON record carries its counts."""
    p = gen("preset", "rmat16", out=tmp_path / "rmat16.edges")
    exp = HASHES["rmat16"]
    out = tmp_path / "rmat16.k"
    args = [os.path.join(BUILD, exe), p] + (["4", str(out)] if exe == "nucleus34" else [str(out)])
    print(f"[cli] {exe} rmat16", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, check=True, timeout=240)
    rec = json.loads(r.stdout)["dataset"]
    assert "error" not in rec, rec
    assert rec["triangle_count"] == exp["T"] and rec["four_clique_count"] == exp["Q"] and rec["K_max"] == exp["K_max"]
    h = hashlib.sha256()
    with open(out, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 24), b""):
            h.update(chunk)
    assert h.hexdigest() == exp["kfile_sha256"]


def _same_as_oracle(g, ref, tag):
    for variant, parts in (("direct", 1), ("ondemand", 1), ("direct", 3)):
        tris, sup, k, st = gpu_run(g, variant, parts)
        assert st["triangles"] == ref["T"] and st["four_cliques"] == ref["Q"], (tag, variant, parts)
        assert np.array_equal(tris, ref["tris"]) and np.array_equal(sup, ref["support"]), (tag, variant, parts)
        assert np.array_equal(k, ref["k"]), (tag, variant, parts)
        if ref["Q"]:
            assert st["k_max"] == ref["K_max"] and st["rounds"] == ref["rounds"], (tag, variant, parts)


def test_random_graph_sweep():
    """1000 small random graphs (the four shapes above, one seed each): every
    run equals the oracle in both variants and with parts > 1."""
    for i in range(1000):
        kind = ("er", "cliques", "hub", "spread")[i % 4]
        g = random_graph(np.random.default_rng(100000 + i), kind)
        ref = oracle34.OracleGraph.from_csr(g.O, g.E, orient=True).run("ondemand", 4)
        _same_as_oracle(g, ref, (kind, 100000 + i))
        if (i + 1) % 200 == 0:
            print(f"[sweep] {i + 1} random graphs equal the oracle", flush=True)


def test_rmat_and_batch_mix_sweep():
    """200 R-MAT graphs with drawn scale (8-12), edge factor (4-32), skew and
    planted cliques, and 100 graphs of the batch mix from drawn seeds: every
    run equals the oracle."""
    rng = np.random.default_rng(7)
    for i in range(300):
        if i < 200:
            sc, ef, a = int(rng.integers(8, 13)), int(rng.integers(4, 33)), float(rng.uniform(0.45, 0.7))
            b = (1 - a) * float(rng.uniform(0.3, 0.45))
            args = (sc, ef, a, b, b, 1000 + i, int(rng.integers(0, 4)), int(rng.integers(5, 16)))
            g = nuc34.gen_rmat(*args)
        else:
            args = (int(rng.integers(10, 1000)), int(rng.integers(0, 256)))
            g = nuc34.gen_batch_graph(*args)
        ref = oracle34.OracleGraph.from_csr(g.O, g.E, orient=True).run("ondemand", 8)
        _same_as_oracle(g, ref, args)
        if (i + 1) % 50 == 0:
            print(f"[sweep] {i + 1} R-MAT / mix graphs equal the oracle (last T={ref['T']} Q={ref['Q']})", flush=True)
