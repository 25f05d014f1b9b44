"""nuc34_decompose_noalloc (every device buffer in the library's per-device
scratch array, no device allocation) against the context path
(nuc34_load_csr + nuc34_decompose + the copies) through the C ABI: bit-exact
triangles, k, T, Q, K_max and rounds, on calls in a row whose later inputs are
smaller than the first (the array keeps the last call's contents)."""
import numpy as np
import pytest

import nuc34

pytestmark = pytest.mark.gpu


def context_run(g, variant):
    with nuc34.Context(0) as ctx:
        ctx.load_graph(g)
        st = ctx.decompose(variant)
        return ctx.copy_triangles(), ctx.copy_k(), st


def test_noalloc_matches_context_path_on_calls_in_a_row():
    graphs = [("rmat12+cliques", nuc34.gen_rmat(12, 16, 0.57, 0.19, 0.19, 21, 2, 14)),
              ("batch-mix", nuc34.gen_batch_graph(4, 5)),
              ("rmat10", nuc34.gen_rmat(10, 8, 0.57, 0.19, 0.19, 22, 0, 0))]
    sizes = [g.m for _, g in graphs]
    assert sizes[0] > sizes[1] and sizes[0] > sizes[2]
    for variant in ("ondemand", "direct"):
        for name, g in graphs:
            print(f"[noalloc] {variant} {name} n={g.n} m={g.m}", flush=True)
            t0, k0, s0 = context_run(g, variant)
            t1, k1, s1 = nuc34.decompose_noalloc(g, variant)
            assert s1["triangles"] == s0["triangles"] > 0 and s1["four_cliques"] == s0["four_cliques"] > 0
            assert s1["k_max"] == s0["k_max"] and s1["rounds"] == s0["rounds"] and s1["unpeeled"] == 0
            assert np.array_equal(t1, t0) and np.array_equal(k1, k0), (variant, name)
            assert 0 < s1["peak_bytes"] <= 256 << 20


def test_noalloc_rejects_what_does_not_fit():
    g = nuc34.gen_batch_graph(4, 5)
    _, _, s0 = context_run(g, "auto")
    with pytest.raises(nuc34.Nuc34Error) as ei:
        nuc34.decompose_noalloc(g, "auto", tri_capacity=s0["triangles"] - 1)
    assert ei.value.code == nuc34.E_INVALID and "tri_capacity" in str(ei.value)
    assert ei.value.stats["triangles"] == s0["triangles"]
    n = 8 << 20  # 96 MiB of input state: over a quarter of the array, refused before any device work
    with pytest.raises(nuc34.Nuc34Error) as ei:
        nuc34.decompose_noalloc((np.zeros(n + 1, np.uint32), np.zeros(0, np.uint32)), "auto")
    assert ei.value.code == nuc34.E_INVALID and "too large" in str(ei.value)
    t1, k1, s1 = nuc34.decompose_noalloc(g, "auto")  # the array is still usable afterwards
    assert s1["triangles"] == s0["triangles"] and len(k1) == s0["triangles"]
