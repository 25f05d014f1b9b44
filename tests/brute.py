"""Independent brute force for tiny graphs (n <= ~60): a textbook restatement of
the (3,4)-nucleus decomposition that shares no code with the oracle or the HIP
path. Used to pin the oracle.

* orientation: keep u->v iff (deg u, u) < (deg v, v), lists sorted by id;
  edge index = position in the concatenated out-lists;
* triangles: every vertex triple u<v<w (rank order) with all three edges, as
  (a, b, c) = (idx(u,w), idx(u,v), idx(v,w)), sorted by (a, b);
* four-cliques: every 4-subset of a triangle's common neighbourhood;
* k: sequential peeling (Sariyuce et al., WWW 2015, Alg. 1): repeatedly remove a
  triangle of minimum current support s, k = max(current level, s), and
  decrement the support of the other three triangles of every K4 of the removed
  triangle whose four triangles were all still present.
"""
from __future__ import annotations

import itertools


def undirected_from_pairs(pairs):
    """Undirected simple graph as from_edges builds it: self-loops dropped,
    n = max id + 1 over the kept pairs (1 if none)."""
    adj = {}
    mx = 0
    for s, d in pairs:
        if s == d:
            continue
        mx = max(mx, s, d)
        adj.setdefault(s, set()).add(d)
        adj.setdefault(d, set()).add(s)
    n = mx + 1
    return n, [sorted(adj.get(u, ())) for u in range(n)]


def decompose(n, nbrs):
    deg = [len(x) for x in nbrs]
    key = lambda v: (deg[v], v)
    out = [[v for v in nbrs[u] if key(u) < key(v)] for u in range(n)]
    idx = {}
    pos = 0
    for u in range(n):
        for v in out[u]:
            idx[(u, v)] = pos
            pos += 1
    adj = [set(x) for x in nbrs]
    tris = []
    for u in range(n):
        for v in out[u]:
            for w in out[u]:
                if w in out[v]:
                    tris.append((idx[(u, w)], idx[(u, v)], idx[(v, w)], frozenset((u, v, w))))
    tris.sort(key=lambda t: (t[0], t[1]))
    tid = {t[3]: i for i, t in enumerate(tris)}
    k4s = []
    seen = set()
    for vs in tid:
        common = set.intersection(*(adj[x] for x in vs))
        for x in common:
            q = frozenset(vs | {x})
            if q in seen:
                continue
            seen.add(q)
            k4s.append([tid[frozenset(c)] for c in itertools.combinations(sorted(q), 3)])
    T = len(tris)
    sup = [0] * T
    by_tri = [[] for _ in range(T)]
    for qi, q in enumerate(k4s):
        for t in q:
            sup[t] += 1
            by_tri[t].append(qi)
    cur = list(sup)
    alive = [True] * T
    k = [0] * T
    level = 0
    for _ in range(T):
        t = min((i for i in range(T) if alive[i]), key=lambda i: (cur[i], i))
        level = max(level, cur[t])
        k[t] = level
        for qi in by_tri[t]:
            q = k4s[qi]
            if all(alive[x] for x in q):
                for x in q:
                    if x != t:
                        cur[x] -= 1
        alive[t] = False
    return {
        "m": pos,
        "tris": [t[:3] for t in tris],
        "T": T,
        "Q": len(k4s),
        "support": sup,
        "k": k,
        "K_max": max(k) if k else 0,
        "out": out,
        "k4s": [frozenset(q) for q in k4s],
    }
