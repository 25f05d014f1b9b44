"""Serving loop (nuc34_serve_*): jobs submitted from several threads come back
bit-exact against their own single-graph decompositions (and the smallest
against the oracle), whichever lane and batch ran them; a batch that exceeds
its lane's budget is split and still returns exact results."""
import threading

import numpy as np
import pytest

import nuc34
import oracle34

pytestmark = pytest.mark.gpu


def single(g, variant="auto"):
    with nuc34.Context(0) as ctx:
        ctx.load_graph(g)
        st = ctx.decompose(variant)
        return ctx.copy_triangles(), ctx.copy_k(), st


def check_jobs(srv, jobs, gs, variant="auto"):
    infos = [srv.wait(j, 120000) for j in jobs]
    assert all(i is not None for i in infos), "job did not finish"
    for j, info, g in zip(jobs, infos, gs):
        assert info["status"] == nuc34.OK, info["error"]
        tris, k = srv.take(j, info["triangles"])
        t1, k1, st1 = single(g, variant)
        assert info["triangles"] == st1["triangles"] and info["four_cliques"] == st1["four_cliques"]
        assert info["t_submit"] <= info["t_start"] <= info["t_done"]
        assert np.array_equal(tris, t1) and np.array_equal(k, k1)
        assert info["k_max"] == (int(k1.max()) if len(k1) else 0)
    return infos


def test_serve_two_lanes_threads():
    print("[serve] two lanes, 3 submitting threads, 30 jobs", flush=True)
    gs = [nuc34.gen_batch_graph(4, i) for i in range(30)]
    large = 150_000  # edge entries: the bigger graphs of the mix go to the large-job lane
    jobs = [None] * len(gs)
    with nuc34.Server(hbm_budget=16 << 30, large_edges=large) as srv:
        def submit(lo):
            for i in range(lo, len(gs), 3):
                jobs[i] = srv.submit(gs[i])
        th = [threading.Thread(target=submit, args=(r,)) for r in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        infos = check_jobs(srv, jobs, gs)
        st = srv.stats()
    assert st["completed"] == len(gs) and st["failed"] == 0
    n_large = sum(g.m >= large for g in gs)
    assert 0 < n_large < len(gs)
    for info, g in zip(infos, gs):
        if g.m >= large:
            assert info["batch_graphs"] == 1
    order = sorted(range(len(gs)), key=lambda i: gs[i].m)
    for i in order[:2]:
        ref = oracle34.OracleGraph.from_csr(gs[i].O, gs[i].E, orient=True).run("host", 4)
        t1, k1, _ = single(gs[i])
        assert np.array_equal(k1, ref["k"]) and np.array_equal(t1, ref["tris"])


def test_serve_batches_and_splits_on_budget():
    print("[serve] forced direct, small-lane budget below the union's footprint", flush=True)
    gs = [nuc34.gen_batch_graph(4, i) for i in range(0, 48, 2)]
    peaks = [single(g, "direct")[2]["peak_bytes"] for g in gs]
    with nuc34.Context(0) as ctx:
        ctx.load_batch(gs)
        union_peak = ctx.decompose("direct")["peak_bytes"]
    if union_peak < 1.2 * max(peaks):
        pytest.skip(f"union footprint {union_peak} too close to the largest single job's {max(peaks)}")
    budget = (max(peaks) + union_peak) // 2  # every job fits alone, the whole batch does not
    # a 2 s batching window gathers all the jobs into the first batch
    with nuc34.Server(hbm_budget=2 * budget, large_fraction=0.5, variant="direct", batch_window_ms=2000) as srv:
        jobs = [srv.submit(g) for g in gs]
        infos = check_jobs(srv, jobs, gs, "direct")
        st = srv.stats()
    print(f"[serve] batches={st['batches']} splits={st['splits']}", flush=True)
    assert st["completed"] == len(gs) and st["failed"] == 0
    assert st["splits"] >= 1 and max(i["batch_graphs"] for i in infos) < len(gs)


def test_serve_empty_and_triangle_free():
    with nuc34.Server(hbm_budget=1 << 30) as srv:
        empty = (np.zeros(1, np.uint32), np.zeros(0, np.uint32))
        path = (np.array([0, 1, 3, 4], np.uint32), np.array([1, 0, 2, 1], np.uint32))
        j0, j1 = srv.submit(empty), srv.submit(path)
        for j in (j0, j1):
            info = srv.wait(j, 60000)
            assert info["status"] == nuc34.OK and info["triangles"] == 0 and info["k_max"] == 0
            tris, k = srv.take(j, 0)
            assert len(tris) == 0 and len(k) == 0
        assert srv.wait(j0, 0) is None  # taken jobs are forgotten


def test_serve_malformed_job_fails_alone():
    """A malformed graph batched with valid ones fails with NUC34_E_INVALID and
    the others still come back exact."""
    print("[serve] one malformed job in a batch of valid ones", flush=True)
    gs = [nuc34.gen_batch_graph(4, i) for i in range(1, 12, 2)]
    bad = (np.array([0, 2, 3, 4], np.uint32), np.array([2, 1, 0, 0], np.uint32))  # vertex 0's list unsorted
    with nuc34.Server(hbm_budget=8 << 30, batch_window_ms=1000) as srv:
        jobs = [srv.submit(g) for g in gs[:3]] + [srv.submit(bad)] + [srv.submit(g) for g in gs[3:]]
        info = srv.wait(jobs[3], 60000)
        assert info is not None and info["status"] == nuc34.E_INVALID, info
        good = jobs[:3] + jobs[4:]
        check_jobs(srv, good, gs)
        st = srv.stats()
    assert st["failed"] == 1 and st["completed"] == len(gs) and st["splits"] >= 1


def test_serve_auto_fit_variant_per_batch():
    """With NUC34_AUTO_FIT each batch picks direct or on demand from its K4
    count and the lane's budget; results are exact either way."""
    print("[serve] auto_fit: small budget lane and large budget lane", flush=True)
    gs = [nuc34.gen_batch_graph(4, i) for i in range(0, 16, 3)]
    with nuc34.Server(hbm_budget=8 << 30, variant="auto_fit") as srv:
        jobs = [srv.submit(g) for g in gs]
        infos = check_jobs(srv, jobs, gs)
    assert {i["variant"] for i in infos} <= {nuc34.DIRECT, nuc34.ONDEMAND}
    assert any(i["variant"] == nuc34.DIRECT for i in infos)


def test_serve_small_jobs_run_beside_a_large_one():
    """Continuous batching: small jobs that arrive while a large job is on the
    device are admitted by their own lane and run beside it (their start and
    end times fall inside the large job's), and every result, the large one's
    included, equals the graph's decomposition alone."""
    print("[serve] R-MAT s16 large job with small jobs arriving beside it", flush=True)
    import time
    big = nuc34.gen_preset("rmat16")
    pool = [nuc34.gen_batch_graph(4, i) for i in range(0, 24, 3)]
    with nuc34.Server(hbm_budget=32 << 30) as srv:
        jb = srv.submit(big)
        jobs, gs = [], []
        while len(jobs) < 40:
            g = pool[len(jobs) % len(pool)]
            jobs.append(srv.submit(g))
            gs.append(g)
            if srv.wait(jb, 0) is not None:
                break
            time.sleep(0.002)
        ib = srv.wait(jb, 120000)
        infos = check_jobs(srv, jobs, gs)
        assert ib is not None and ib["status"] == nuc34.OK, ib
        tb, kb = srv.take(jb, ib["triangles"])
    inside = [i for i in infos if ib["t_start"] <= i["t_start"] and i["t_done"] <= ib["t_done"]]
    print(f"[serve] {len(inside)} of {len(infos)} small jobs ran inside the large job's "
          f"{1e3 * (ib['t_done'] - ib['t_start']):.0f} ms", flush=True)
    assert inside
    t1, k1, _ = single(big)
    assert np.array_equal(tb, t1) and np.array_equal(kb, k1)


def test_serve_small_jobs_overtake_a_long_small_batch():
    """Continuous batching inside the small-job lane: with two small-job
    workers, small jobs that arrive while a long small-lane batch is on the
    device start on the other worker and finish before that batch ends; with
    one worker they wait for it. Every result, alone or in the mix, equals the
    graph's own decomposition."""
    import time
    big = nuc34.gen_preset("rmat16")  # lane thresholds raised below, so it is a small-lane batch of one
    pool = [nuc34.gen_batch_graph(4, i) for i in range(0, 24, 3)]
    for workers in (2, 1):
        print(f"[serve] long small-lane batch, {workers} small-job worker(s)", flush=True)
        with nuc34.Server(hbm_budget=32 << 30, large_edges=1 << 40, large_wedges=1 << 62,
                          small_workers=workers) as srv:
            jb = srv.submit(big)
            time.sleep(0.005)  # the big job is taken alone
            jobs, gs = [], []
            while len(jobs) < 40:
                g = pool[len(jobs) % len(pool)]
                jobs.append(srv.submit(g))
                gs.append(g)
                if srv.wait(jb, 0) is not None:
                    break
                time.sleep(0.002)
            ib = srv.wait(jb, 120000)
            infos = check_jobs(srv, jobs, gs)
            assert ib is not None and ib["status"] == nuc34.OK, ib
            tb, kb = srv.take(jb, ib["triangles"])
        assert ib["batch_graphs"] == 1
        inside = [i for i in infos if ib["t_start"] <= i["t_start"] and i["t_done"] <= ib["t_done"]]
        print(f"[serve] {len(inside)} of {len(infos)} small jobs started and finished inside the "
              f"{1e3 * (ib['t_done'] - ib['t_start']):.0f} ms batch", flush=True)
        if workers == 2:
            assert inside and all(i["batch_id"] != ib["batch_id"] for i in infos)
        else:
            assert all(i["t_start"] >= ib["t_done"] for i in infos)
        t1, k1, _ = single(big)
        assert np.array_equal(tb, t1) and np.array_equal(kb, k1)


def test_serve_poisson_stream_with_rmat18_jobs():
    """The serving workload: a Poisson stream (200 jobs/s, seed 4, 1 s of
    arrivals) of the many-graphs mix plus R-MAT scale 18 jobs at a 5% share,
    under HBM admission with the variant chosen per batch from its footprint
    (NUC34_AUTO_FIT: the scale-18 lists, 114 GB, do not fit a large worker's
    share of the budget; small batches' do). Every R-MAT scale-18 result
    hashes to the golden raw k and triangle arrays, every small job equals its
    own decomposition, and the smallest ones equal the oracle's."""
    import hashlib
    import json
    import os
    import queue
    import time
    from conftest import GOLDEN
    gold = json.load(open(os.path.join(GOLDEN, "hashes.json")))["rmat18"]
    rng = np.random.default_rng(4)
    pool = [nuc34.gen_batch_graph(4, i) for i in range(64)]
    large = nuc34.gen_preset("rmat18")
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1 / 200.0)
        if t > 1.0:
            break
        big = bool(rng.random() < 0.05)
        arrivals.append((t, big, int(rng.integers(len(pool)))))
    n_large = sum(b for _, b, _ in arrivals)
    print(f"[serve] Poisson stream: {len(arrivals)} jobs, {n_large} R-MAT s18", flush=True)
    assert n_large >= 5
    alone = {}
    for i in sorted({i for _, b, i in arrivals if not b}):
        t1, k1, _ = single(pool[i])
        alone[i] = (t1, k1)
    bad, done, variants = [], [], set()
    q = queue.Queue()
    with nuc34.Server(0, variant="auto_fit") as srv:
        import threading

        def collect():
            while True:
                item = q.get()
                if item is None:
                    return
                job, big, i = item
                info = srv.wait(job, 600000)
                if info is None or info["status"] != nuc34.OK:
                    bad.append((big, i, info))
                    continue
                tris, k = srv.take(job, info["triangles"])
                variants.add((big, info["variant"]))
                if big:
                    ok = (hashlib.sha256(k.tobytes()).hexdigest() == gold["k_raw_sha256"] and
                          hashlib.sha256(tris.tobytes()).hexdigest() == gold["tri_raw_sha256"] and
                          info["k_max"] == gold["K_max"] and info["four_cliques"] == gold["Q"])
                    print(f"[serve] R-MAT s18 job done, {1e3 * (info['t_done'] - info['t_submit']):.0f} ms, "
                          f"bit-exact={ok}", flush=True)
                else:
                    ok = np.array_equal(tris, alone[i][0]) and np.array_equal(k, alone[i][1])
                if not ok:
                    bad.append((big, i, "mismatch"))
                done.append((big, info["t_done"] - info["t_submit"]))

        th = threading.Thread(target=collect)
        th.start()
        t0 = srv.now()
        for ta, big, i in arrivals:
            d = t0 + ta - srv.now()
            if d > 0:
                time.sleep(d)
            q.put((srv.submit(large if big else pool[i]), big, i))
        q.put(None)
        th.join()
        st = srv.stats()
    assert not bad, bad[:5]
    assert len(done) == len(arrivals) and st["completed"] == len(arrivals) and st["failed"] == 0
    assert (True, nuc34.ONDEMAND) in variants and (False, nuc34.DIRECT) in variants
    lat = np.array([d for b, d in done if not b])
    print(f"[serve] small p50/p99 {1e3 * np.percentile(lat, 50):.1f}/{1e3 * np.percentile(lat, 99):.1f} ms", flush=True)
    order = sorted(alone, key=lambda i: pool[i].m)
    for i in order[:2]:
        ref = oracle34.OracleGraph.from_csr(pool[i].O, pool[i].E, orient=True).run("host", 4)
        assert np.array_equal(alone[i][1], ref["k"]) and np.array_equal(alone[i][0], ref["tris"])
