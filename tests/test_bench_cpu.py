"""bench.py's host-side pieces on the CPU: the refusal of more than one GPU, and
the algorithmic byte model against counts known in closed form on cliques."""
import os
import subprocess
import sys
from math import comb

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "measures one GPU" in r.stderr


def _clique(n):
    """K_n oriented from lower to higher id."""
    O = np.zeros(n + 1, dtype=np.uint32)
    E = []
    for u in range(n):
        E += list(range(u + 1, n))
        O[u + 1] = len(E)
    return O, np.array(E, dtype=np.uint32)


@pytest.mark.parametrize("n", [4, 5, 7])
def test_graph_stats_on_cliques(n):
    # d+(u) = n-1-u; the length-2 paths u < v < w are the triangles, C(n, 3)
    O, E = _clique(n)
    g = bench.graph_stats(O, E)
    assert (g["n"], g["m"], g["P"]) == (n, comb(n, 2), comb(n, 3))
    assert g["sum_dplus2"] == (n - 1) * n * (2 * n - 1) // 6


def test_byte_model_on_k4():
    # K4: T = 4, Q = 1, every edge on l = 2 triangles (sum_l2 = 6 * 4), and the
    # peel retires all four triangles in round 0 with no survivor to debit (D = 0)
    O, E = _clique(4)
    g = bench.graph_stats(O, E)
    g["sum_l2"] = 24
    st = {"triangles": 4, "four_cliques": 1, "debits": 0, "variant": 1}
    # lists: 4 (sum d+^2 = 14, P = 4) = 72; CSR: 8 * 5 + 4 * 6 = 64
    assert bench.algorithmic_bytes("tri_enum_bits", g, st) == 136
    assert bench.algorithmic_bytes("tri_enum_emit", g, st) == 136 + 48
    # sum_t (12 + 8 L(t)) with L(t) = 6: 4 * 60 = 240; + 4 T
    assert bench.algorithmic_bytes("support_bits", g, st) == 256
    assert bench.algorithmic_bytes("peel_rounds", g, st) == 256
    st["variant"] = 0  # direct: ids 8 (T + 1), K4 records and incidence 64 Q, 4 T
    assert bench.algorithmic_bytes("peel_rounds", g, st) == 40 + 64 + 16
    assert bench.algorithmic_bytes("k4_lists", g, st) == 256 + 48
    assert bench.algorithmic_bytes("no such kernel", g, st) == 0


def test_atomics_summary():
    st = {"debits": 400, "ms_peel": 2.0}
    a = bench.peel_atomics(st, {"atomic_gops": 1e-4}, {"atomic_requests_per_step": 100})
    assert a["debits_per_request"] == 4 and a["counted"] == "requests"
    assert a["achieved"] == pytest.approx(100 / 2e-3 / 1e9) and a["floor_ms"] == pytest.approx(1.0)
    assert bench.peel_atomics(st, None, None)["counted"] == "decrements"
    assert bench.peel_atomics({"debits": 0, "ms_peel": 1.0}, None, None) is None
