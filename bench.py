#!/usr/bin/env python3
"""Headline benchmark: (3,4)-nucleus decomposition on one MI355X.

Workload (default): R-MAT scale 20, edge factor 16, (a,b,c,d) = (.57,.19,.19,.05),
seed 2, plus 64 planted 30-cliques, symmetrised and deduplicated (424 M triangles,
17.0 G four-cliques), on-demand variant: the largest single-GPU configuration
(the direct variant's K4 lists would need 816 GB).

A step = one decomposition from the oriented CSR resident in HBM to k for every
triangle resident in HBM: triangle enumeration + bit-packed incidence, K4
support, peel. Between steps the context drops the derived stages (buffers
return to its reuse cache, so the timed steps allocate nothing).
value = four-cliques / second = Q / (mean step time), printed only when the raw
k and triangle arrays of the last step hash to the golden values (the
reference's on-demand algorithm on the same graph) and T, Q, K_max and the
round count agree; otherwise value is null and the exit status is 1.

Prints one JSON line with the driver's contract, plus:
  roofline      the dominant kernel group (the peel rounds): algorithmic bytes
                per launched round (the per-unit formulas of the hot-path survey,
                with L(t) summed from the per-edge triangle counts) over its
                average round time from HIP events on the context stream;
                traffic = its HBM bytes per round from two rocprofv3 --pmc
                child runs (FETCH_SIZE, WRITE_SIZE) scaled by a copy of known
                size in the same process;
  roofline.atomics  the peel's atomic line requests (a third --pmc child run)
                per second against the measured random-atomic rate of the
                device (one line per lane), and the decrements per request;
  cpu_baseline  the reference's own host path (oracle/_ref/ref34, compiled from
                the reference headers) on R-MAT scale 16 (the same generator at
                a size the host path holds: about 64 bytes per four-clique), timed
                on this host's cores beside the GPU's time on the same file; and
                the C++ restatement of the reference's on-demand algorithm timed
                once on the full default workload (profiles/, read if present);
  batched       512 graphs of the seeded mixed batch decomposed as one disjoint
                union versus one decomposition per graph (graphs/s, bit-exact);
  device_rates  measured copy GB/s and random 4-byte atomic / load rates;
  config.kernel_gbs      algorithmic GB/s of every timed kernel group, by the
                survey's per-unit formulas, which price the reference's merge over
                incidence lists of length l(e): where this build reads bit rows
                instead (the support count) the model rate can pass the HBM peak;
  config.kernel_physical_gbs  measured HBM GB/s of the same groups (PMC fetch +
                write bytes of the profiled step over the timed kernel time);
  config.end_to_end_ms   parse, load (host to device, checks, orientation),
                decomposition and the copy of k and the triangles back.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gpu-nucleus34_amd")
GEN = os.path.join(PKG, "_build", "nuc34_gen")
HASHES = os.path.join(ROOT, "tests", "golden", "hashes.json")
RESTATEMENT_RECORD = os.path.join(ROOT, "profiles", "rmat20k30_cpu_ondemand.json")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

WORKLOADS = {
    "rmat20k30": {"gen": ["preset", "rmat20k30"], "variant": "ondemand", "parts": 1,
                  "desc": "R-MAT scale 20, edge factor 16, (a,b,c,d)=(.57,.19,.19,.05), seed 2, + 64 planted K30, "
                          "on-demand variant"},
    "rmat16": {"gen": ["preset", "rmat16"], "variant": "direct", "parts": 1,
               "desc": "R-MAT scale 16, edge factor 16, (a,b,c,d)=(.57,.19,.19,.05), seed 1, direct variant"},
    "rmat16-parts4": {"gen": ["preset", "rmat16"], "variant": "direct", "parts": 4,
                      "desc": "R-MAT scale 16, edge factor 16, seed 1, direct variant, parts=4"},
    "rmat16-ondemand": {"gen": ["preset", "rmat16"], "variant": "ondemand", "parts": 1,
                        "desc": "R-MAT scale 16, edge factor 16, seed 1, on-demand variant"},
    "er1000": {"gen": ["preset", "er1000"], "variant": "direct", "parts": 1,
               "desc": "Erdos-Renyi G(1000, 0.1), seed 0, direct variant"},
    "rmat18": {"gen": ["preset", "rmat18"], "variant": "direct", "parts": 1,
               "desc": "R-MAT scale 18, edge factor 16, seed 1, direct variant (the serving mix's large job)"},
    "rmat18-ondemand": {"gen": ["preset", "rmat18"], "variant": "ondemand", "parts": 1,
                        "desc": "R-MAT scale 18, edge factor 16, seed 1, on-demand variant"},
}

# Launch shapes chosen by nuc34_autotune on one MI355X (bench.py --autotune
# prints the tuner's choice); shapes change speed, never results.
TUNED = {
    "rmat20k30": {"peel_debit_grid": 8192, "peel_tile_cap": 8, "peel_rounds_per_launch": 64, "tri_tile_edges": 16,
                  "peel_finish_grid": 1024, "tri_schedule": 1},
    "rmat16": {"peel_debit_grid": 2048, "peel_tile_cap": 16, "peel_rounds_per_launch": 32, "k4_list_grid_cap": 8192,
               "tri_tile_edges": 64, "peel_finish_grid": 1024, "tri_schedule": 1},
    "batch": {"peel_debit_grid": 2048, "peel_tile_cap": 32, "peel_rounds_per_launch": 16, "k4_list_grid_cap": 32768,
              "tri_tile_edges": 64, "peel_finish_grid": 512, "batch_peel": 2, "tri_schedule": 0},
}


def tuned_for(workload: str) -> dict | None:
    return TUNED.get(workload.split("-")[0] if workload.startswith("rmat16") else workload)


def graph_stats(O: np.ndarray, E: np.ndarray) -> dict:
    d = np.diff(O.astype(np.int64))
    P = int(d[E].sum()) if len(E) else 0  # length-2 paths: sum over edges of d+(head)
    return {"n": len(d), "m": len(E), "P": P, "sum_dplus2": int((d * d).sum())}


def algorithmic_bytes(kernel: str, g: dict, st: dict) -> float:
    """Bytes per launch of `kernel` by the per-unit formulas of the hot-path
    survey: n, m, T, Q; D = support decrements applied; L(t) = l(a)+l(b)+l(c)
    with l(e) the triangles on edge e, so sum_t L(t) = sum_e l(e)^2 (g["sum_l2"]);
    the merge intersection reads 4 bytes per entry of N+(u) and N+(v) for every
    oriented edge (u,v): sum_u d+(u)^2 + P."""
    T, Q, m, n, D = st["triangles"], st["four_cliques"], g["m"], g["n"], st["debits"]
    lists = 4 * (g["sum_dplus2"] + g["P"])
    enum_bytes = 12 * T + 8 * g["sum_l2"]  # sum_t (12 + 8 L(t))
    if kernel == "tri_enum_bits":  # pass 1 of 2: list reads, CSR; the triangle records come in pass 2
        return lists + 8 * (n + 1) + 4 * m
    if kernel == "tri_enum_emit":
        return lists + 12 * T + 8 * (n + 1) + 4 * m
    if kernel == "support_bits":
        return enum_bytes + 4 * T
    if kernel == "k4_lists":  # direct: the K4 records and the triangle -> K4 incidence written
        return enum_bytes + 4 * T + 16 * Q + 32 * Q
    if kernel == "peel_rounds":
        if st["variant"] == 0:  # direct
            return 8 * (T + 1) + 16 * Q + 32 * Q + 16 * Q + 8 * D + 4 * T
        return enum_bytes + 8 * D + 4 * T  # on demand: every triangle re-enumerates its K4s once, at death
    return 0.0


# Kernels that make up each timed group (the per-kernel event timers of the
# library time these groups; the peel group is one captured graph of rounds).
GROUP_KERNELS = {
    "tri_enum_bits": ("k_tri_enum<false>", "k_tri_vertex_wave<false>", "k_tri_vertex_group<false"),
    "tri_enum_emit": ("k_tri_enum<true>", "k_tri_vertex_wave<true>", "k_tri_vertex_group<true"),
    "support_bits": ("k_support_bits",),
    "k4_lists": ("k_k4_lists",),
    "peel_rounds": ("k_peel_collect", "k_peel_debit", "k_peel_finish", "k_peel_graphs"),
}


def _pmc_rows(db_path: str):
    import sqlite3
    con = sqlite3.connect(db_path)
    rows = con.execute("select name, counter_value from pmc_events order by dispatch_id").fetchall()
    con.close()
    return [(n.replace("(anonymous namespace)::", "").split("(")[0], v) for n, v in rows]


def calibration_copy(mib: int) -> None:
    """One device-to-device copy of `mib` MiB through the HIP runtime (no torch:
    with torch loaded, the profiler's counter mode breaks HIP event calls)."""
    import ctypes as C
    path = "/opt/rocm/lib/libamdhip64.so"
    hip = C.CDLL(path if os.path.exists(path) else "libamdhip64.so")
    n = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    for call in (lambda: hip.hipMalloc(C.byref(a), n), lambda: hip.hipMalloc(C.byref(b), n),
                 lambda: hip.hipMemset(a, 1, n), lambda: hip.hipMemcpy(b, a, n, 3),  # 3: device to device
                 lambda: hip.hipDeviceSynchronize(), lambda: hip.hipFree(a), lambda: hip.hipFree(b)):
        rc = call()
        if rc != 0:
            raise RuntimeError(f"HIP error {rc} in the calibration copy")


# Returning atomic requests from the CUs to L2, one per 64-byte line per wave
# instruction: what a device atomic costs, however many lanes share the line.
ATOMIC_REQ = "TCP_TCC_ATOMIC_WITH_RET_REQ_sum"


def run_traffic(workload: str, edges: str, group: str, launches: float, tmp: str, tuning: dict | None,
                calib_mib: int = 1024) -> dict | None:
    """HBM bytes per launch of `group` from two rocprofv3 --pmc runs of a
    one-step bench on the same edge file (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch), each scaled by what the same counter reads for a device-to-device
    copy of `calib_mib` MiB made first in the same process, and the group's
    returning atomic requests per step from a third run (ATOMIC_REQ). The child
    runs with the launch shapes of the timed run (`tuning`). None when rocprofv3
    is unavailable or fails."""
    prof = "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    child = [sys.executable, os.path.abspath(__file__), "--steps", "1", "--warmup", "0", "--no-batch",
             "--no-cpu-baseline", "--no-traffic", "--no-rates", "--no-other-variant", "--no-verify",
             "--workload", workload, "--edges", edges,
             *(["--tuning", json.dumps(tuning)] if tuning else []), "--copy-calibration-mib", str(calib_mib)]
    res = {}
    # Peel rounds launched one kernel at a time instead of as captured graphs:
    # the same kernels with the same launch shapes, so the per-dispatch counters
    # are unchanged. A --kernel-trace run of this library faulted inside
    # hipGraphLaunch, in the profiler's own code called from the HSA runtime,
    # so the profiled children keep away from graph launches.
    env = dict(os.environ, NUC34_NO_GRAPH="1")
    for counter in ("FETCH_SIZE", "WRITE_SIZE", ATOMIC_REQ):
        d = os.path.join(tmp, counter)
        print(f"[bench] rocprofv3 --pmc {counter} child run", file=sys.stderr, flush=True)
        try:
            # the profiler can exit nonzero after writing complete results, so
            # the database decides, not the exit status
            rc = subprocess.run([prof, "--pmc", counter, "-d", d, "-o", "run", "--", *child], timeout=240,
                                stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, env=env).returncode
        except (subprocess.SubprocessError, OSError) as e:
            print(f"traffic: {counter} run failed: {e}", file=sys.stderr)
            return None
        dbs = [os.path.join(r, f) for r, _, fs in os.walk(d) for f in fs if f.endswith("results.db")]
        if not dbs:
            print(f"traffic: {counter} run wrote no database (exit {rc})", file=sys.stderr)
            return None
        try:
            rows = _pmc_rows(dbs[0])
        except Exception as e:  # noqa: BLE001 - report and fall back to null
            print(f"traffic: {counter} database unreadable (exit {rc}): {e}", file=sys.stderr)
            return None
        if counter == ATOMIC_REQ:  # a count, not bytes: no calibration
            res[counter] = sum(v for n, v in rows if any(k in n for k in GROUP_KERNELS[group]))
            continue
        copies = [v for n, v in rows if "copyBuffer" in n]
        if not copies:
            print(f"traffic: {counter} database has no calibration copy (exit {rc}, {len(rows)} rows)", file=sys.stderr)
            return None
        scale = calib_mib * 1024.0 / max(copies)  # counter KiB -> true KiB
        per_group = {g: sum(v for n, v in rows if any(k in n for k in names)) * scale * 1024.0
                     for g, names in GROUP_KERNELS.items()}  # bytes in the one profiled step
        res[counter] = (per_group[group] / launches, scale, per_group)
    return {"bytes_per_launch": res["FETCH_SIZE"][0] + res["WRITE_SIZE"][0],
            "fetch_bytes_per_launch": res["FETCH_SIZE"][0], "write_bytes_per_launch": res["WRITE_SIZE"][0],
            "fetch_scale": res["FETCH_SIZE"][1], "write_scale": res["WRITE_SIZE"][1],
            "group_bytes_per_step": {g: res["FETCH_SIZE"][2][g] + res["WRITE_SIZE"][2][g]
                                     for g in GROUP_KERNELS if res["FETCH_SIZE"][2][g] + res["WRITE_SIZE"][2][g] > 0},
            "calibration": f"{calib_mib} MiB device-to-device copy in the profiled process",
            "atomic_requests_per_step": res[ATOMIC_REQ]}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def sha256_of(a: np.ndarray) -> str:
    h = hashlib.sha256()
    mv = memoryview(np.ascontiguousarray(a)).cast("B")
    for i in range(0, len(mv), 1 << 28):
        h.update(mv[i:i + (1 << 28)])
    return h.hexdigest()


def verify(preset: str, st: dict, k: np.ndarray, tris: np.ndarray) -> dict:
    """The last step's results against the golden values for this graph: T, Q,
    K_max, rounds and the sha256 of the raw k and (a,b,c) arrays."""
    gold = json.load(open(HASHES)).get(preset)
    if not gold or "k_raw_sha256" not in gold:
        return {"checked": False, "bit_exact": None}
    got = {"T": st["triangles"], "Q": st["four_cliques"], "K_max": st["k_max"], "rounds": st["rounds"],
           "k_raw_sha256": sha256_of(k), "tri_raw_sha256": sha256_of(tris)}
    bad = [f for f in got if f in gold and got[f] != gold[f]]
    return {"checked": True, "bit_exact": not bad, "mismatched": bad, "k_raw_sha256": got["k_raw_sha256"],
            "tri_raw_sha256": got["tri_raw_sha256"],
            "against": "tests/golden/hashes.json (the oracle's on-demand and host restatements; for er1000 and "
                       "rmat16 also the reference's own host path)"}


def run_same_input_pair(nuc34, tmp: str, threads: int) -> dict:
    """R-MAT scale 16 (seed 1): the reference's host path and the GPU on the same
    .edges file in the same run, each checked against the golden hashes."""
    path = os.path.join(tmp, "rmat16.edges")
    subprocess.run([GEN, "preset", "rmat16", path], check=True, stdout=subprocess.DEVNULL)
    g = nuc34.load_dataset(path)
    gpu = {}
    with nuc34.Context(0) as ctx:
        ctx.load_graph(g)
        ctx.set_tuning(**TUNED["rmat16"])
        for variant in ("direct", "ondemand"):
            ctx.reset()
            ctx.decompose(variant)
            t0 = time.perf_counter()
            for _ in range(3):
                ctx.reset()
                st = ctx.decompose(variant)
            ms = (time.perf_counter() - t0) / 3 * 1e3
            ver = verify("rmat16", st, ctx.copy_k(), ctx.copy_triangles())
            gpu[variant] = {"ms": ms, "value": st["four_cliques"] / (ms / 1e3), "bit_exact": ver["bit_exact"]}
    ref = os.path.join(ROOT, "oracle", "_ref", "ref34")
    kind = "reference"
    kfile = os.path.join(tmp, "rmat16.ref.k")
    if os.path.exists(ref):
        args = [ref, path, kfile, str(threads)]
    else:  # the reference tree was not there to build it: the oracle's restatement of the same path
        ref = os.path.join(ROOT, "oracle", "_build", "oracle34")
        kind = "port"
        args = [ref, path, "--parts", str(threads), "--kfile", kfile]
    print(f"[bench] CPU host path ({kind}) on R-MAT scale 16, {threads} threads", file=sys.stderr, flush=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run(args, check=True, capture_output=True, text=True, env=env).stdout
    rec = json.loads(out.strip().splitlines()[-1])
    secs = sum(rec.get(k, 0.0) for k in ("tri_s", "tri_graph_s", "qud_s", "qud_graph_s", "peel_s"))
    h = hashlib.sha256()
    with open(kfile, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 24), b""):
            h.update(chunk)
    gold = json.load(open(HASHES))["rmat16"]
    cpu_value = rec["Q"] / secs
    return {"value": cpu_value, "unit": "4-cliques/s", "cores": threads, "kind": kind, "cpu": cpu_model(),
            "threads_note": "OpenMP loops on `cores` threads; std::execution::par runs serial (libstdc++ without TBB)",
            "phases_s": {k: rec.get(k, 0.0) for k in ("tri_s", "tri_graph_s", "qud_s", "qud_graph_s", "peel_s")},
            "kfile_matches_golden": h.hexdigest() == gold["kfile_sha256"],
            "sample": f"R-MAT scale 16, edge factor 16, same (a,b,c,d), seed 1 (T={rec['T']}, Q={rec['Q']}, "
                      f"K_max={rec['K_max']}): the default workload's generator at a size the reference's host path "
                      f"holds in memory; decomposition {secs:.2f} s, load excluded",
            "same_input": {"file": "R-MAT scale 16 seed 1 (.edges)", "cpu_s": secs,
                           "gpu": gpu, "speedup_direct": gpu["direct"]["value"] / cpu_value,
                           "speedup_ondemand": gpu["ondemand"]["value"] / cpu_value}}


def run_batched(nuc34, count: int, seed: int = 3, autotune: bool = False) -> dict:
    """Many-graphs workload: `count` graphs of the seeded mixed batch, decomposed
    as one disjoint union (one pass of every kernel, shared peel rounds) versus one
    decomposition per graph; both include upload, device orientation and the
    copy of k back to the host. Every graph's slice of the batched result must
    equal its own decomposition."""
    gs = [nuc34.gen_batch_graph(seed, i) for i in range(count)]
    with nuc34.Context(0) as ctx:
        ctx.load_batch(gs)  # launch shapes for the union, used by both modes
        if autotune:
            tuning, _ = ctx.autotune("auto")
        else:
            ctx.set_tuning(**TUNED["batch"])
            tuning = ctx.tuning()
        ctx.load_batch(gs[:8])  # warm-up of both paths (allocator cache, code objects)
        ctx.decompose("auto")
        ctx.load_graph(gs[0])
        ctx.decompose("auto")
        t0 = time.perf_counter()
        ctx.load_batch(gs)
        stb = ctx.decompose("auto")
        k_b = ctx.copy_k()  # results back, as the per-graph mode does below
        t1 = time.perf_counter()
        infos = [ctx.batch_graph(g) for g in range(count)]
        t2 = time.perf_counter()
        ks = []
        for g in gs:
            ctx.load_graph(g)
            ctx.decompose("auto")
            ks.append(ctx.copy_k())
        t3 = time.perf_counter()
    plain_s = t3 - t2
    exact = all(np.array_equal(k_b[i["tri_first"]:i["tri_first"] + i["tri_count"]], kk) for i, kk in zip(infos, ks))
    return {"graphs": count, "seed": seed, "vertices": int(sum(g.n for g in gs)), "edges": int(sum(g.m for g in gs) // 2),
            "triangles": stb["triangles"], "four_cliques": stb["four_cliques"], "rounds": stb["rounds"],
            "batched_s": t1 - t0, "plain_s": plain_s, "graphs_per_s_batched": count / (t1 - t0),
            "graphs_per_s_plain": count / plain_s, "speedup": plain_s / (t1 - t0), "bit_exact_vs_plain": bool(exact),
            "tuning": tuning}


def run_serving(nuc34, seconds: float, rate: float, large_share: float, seed: int = 4, pool: int = 256,
                small_workers: int = 0) -> dict:
    """Serving workload: Poisson arrivals at `rate` jobs/s for `seconds`, seed
    `seed`; each job is a graph of the seeded mixed batch (pool of `pool`
    graphs) or, with probability `large_share`, an R-MAT scale 18 edge factor 16
    graph (two seeds). Jobs go to a nuc34 server in real time; a collector
    thread takes every result and checks it bit-exact against the same graph's
    own decomposition. Latency = result on the host minus submission."""
    import queue
    import threading

    rng = np.random.default_rng(seed)
    small = [nuc34.gen_batch_graph(seed, i) for i in range(pool)]
    large = [nuc34.gen_rmat(18, 16, 0.57, 0.19, 0.19, s_, 0, 0) for s_ in (1, 2)]
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            break
        big = bool(rng.random() < large_share)
        arrivals.append((t, big, int(rng.integers(len(large) if big else len(small)))))
    used = sorted({(b, i) for _, b, i in arrivals})
    expect = {}
    with nuc34.Context(0) as ctx:  # reference results: each graph decomposed on its own
        for b, i in used:
            ctx.load_graph(large[i] if b else small[i])
            ctx.decompose("auto")
            expect[(b, i)] = (ctx.copy_triangles(), ctx.copy_k())
    results, bad = [], []
    q = queue.Queue()
    with nuc34.Server(0, small_workers=small_workers) as srv:
        def collect():
            while True:
                item = q.get()
                if item is None:
                    return
                job, key = item
                info = srv.wait(job, -1)
                tris, k = srv.take(job, info["triangles"])
                et, ek = expect[key]
                if info["status"] != nuc34.OK or not (np.array_equal(k, ek) and np.array_equal(tris, et)):
                    bad.append((key, info["status"], info["error"]))
                results.append((key[0], info["t_submit"], info["t_done"], info["batch_graphs"]))
                if len(results) % 1000 == 0:
                    print(f"[serve] {len(results)} of {len(arrivals)} jobs collected", file=sys.stderr, flush=True)
        th = threading.Thread(target=collect)
        th.start()
        t0 = srv.now()
        for ta, b, i in arrivals:
            d = t0 + ta - srv.now()
            if d > 0:
                time.sleep(d)
            q.put((srv.submit(large[i] if b else small[i]), (b, i)))
        q.put(None)
        th.join()
        st = srv.stats()
    lat = {c: np.array([r[2] - r[1] for r in results if r[0] == c]) for c in (False, True)}
    allat = np.array([r[2] - r[1] for r in results])
    span = max(r[2] for r in results) - min(r[1] for r in results)
    pct = lambda a, p: float(np.percentile(a, p)) * 1e3 if len(a) else None
    return {"seconds": seconds, "rate": rate, "large_share": large_share, "seed": seed, "jobs": len(results),
            "large_jobs": int(sum(r[0] for r in results)), "bit_exact": not bad, "mismatches": bad[:5],
            "graphs_per_s": len(results) / span, "span_s": span,
            "p50_ms": pct(allat, 50), "p99_ms": pct(allat, 99),
            "small_p50_ms": pct(lat[False], 50), "small_p99_ms": pct(lat[False], 99),
            "large_p50_ms": pct(lat[True], 50), "large_p99_ms": pct(lat[True], 99),
            "mean_batch_graphs": float(np.mean([r[3] for r in results if not r[0]])) if lat[False].size else None,
            "server": st}


def peel_atomics(st: dict, rates: dict | None, traffic: dict | None) -> dict | None:
    """The peel's support decrements are device-scope atomics, and a device
    atomic costs one request per 64-byte line per wave instruction: with the
    peel rounds' request count (PMC) the rate and floor are in requests against
    the device's measured random-atomic rate (one line per lane); without it, in
    decrements (an upper bound on the requests)."""
    if not st.get("debits") or not st.get("ms_peel"):
        return None
    sec = st["ms_peel"] / 1e3
    peak = rates["atomic_gops"] if rates else None
    req = traffic.get("atomic_requests_per_step") if traffic else None
    n = req if req else st["debits"]
    got = n / sec / 1e9
    return {"debits_per_step": st["debits"], "requests_per_step": req,
            "debits_per_request": st["debits"] / req if req else None,
            "counted": "requests" if req else "decrements",
            "achieved": got, "unit": "G requests/s" if req else "G atomics/s", "peak_measured": peak,
            "frac": got / peak if peak else None,
            "floor_ms": n / (peak * 1e9) * 1e3 if peak else None}


def restatement_record() -> dict | None:
    """The C++ restatement of the reference's on-demand algorithm timed once on
    the default workload (R-MAT s20 + 64 K30) on the host cores of the MI355X
    machine (profiles/), if present."""
    try:
        return json.load(open(RESTATEMENT_RECORD))
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat20k30", choices=sorted(WORKLOADS))
    ap.add_argument("--edges", default=None, help="use this .edges file instead of generating the workload's")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-graph batched-vs-plain comparison")
    ap.add_argument("--batch-graphs", type=int, default=512)
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 --pmc runs for roofline.traffic")
    ap.add_argument("--autotune", action="store_true", help="tune launch shapes on the graph instead of the committed ones")
    ap.add_argument("--tuning", default=None,
                    help="launch shapes as a JSON object of nuc34_tuning fields (default: the committed shapes)")
    ap.add_argument("--no-other-variant", action="store_true",
                    help="skip timing the other variant (direct <-> on demand) on the same graph")
    ap.add_argument("--no-rates", action="store_true",
                    help="skip measuring the device's copy bandwidth and random atomic / load rates")
    ap.add_argument("--no-verify", action="store_true", help="skip the golden-hash check (profiler child runs)")
    ap.add_argument("--serve-seconds", type=float, default=0.0,
                    help="also run the serving workload (Poisson arrivals) for this many seconds of arrivals")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-large-share", type=float, default=0.05)
    ap.add_argument("--serve-small-workers", type=int, default=0)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first make a device-to-device copy of this many MiB (counter calibration under --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: batches and served graphs "
              "would shard by graph and one large graph by vertex range of its lowest vertex; the library "
              "runs on one device", file=sys.stderr)
        sys.exit(2)
    import nuc34

    if args.copy_calibration_mib:
        calibration_copy(args.copy_calibration_mib)
    wl = WORKLOADS[args.workload]
    preset = wl["gen"][-1]
    with tempfile.TemporaryDirectory() as tmp:
        path = args.edges or os.path.join(tmp, "graph.edges")
        if not args.edges:
            subprocess.run([GEN, *wl["gen"], path], check=True, stdout=subprocess.DEVNULL)
        t_parse = time.perf_counter()
        g = nuc34.load_dataset(path)
        t_parse = time.perf_counter() - t_parse
        ctx = nuc34.Context(0)
        t_load = time.perf_counter()
        ctx.load_graph(g)  # host to device, checks, orientation on the device
        t_load = time.perf_counter() - t_load
        del g
        O, E = ctx.oriented_csr()
        gs = graph_stats(O, E)
        del O, E
        if args.autotune:
            tuning = ctx.autotune(wl["variant"])[0]
        else:
            shapes = json.loads(args.tuning) if args.tuning else tuned_for(args.workload)
            if shapes:
                ctx.set_tuning(**shapes)
            tuning = ctx.tuning()
        print(f"[bench] {args.workload}: {args.warmup} warm-up + {args.steps} timed steps", file=sys.stderr, flush=True)
        for _ in range(args.warmup):
            ctx.reset()
            ctx.decompose(wl["variant"], wl["parts"])
        steps = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.reset()
            steps.append(ctx.decompose(wl["variant"], wl["parts"]))
        t1 = time.perf_counter()
        st = steps[-1]
        t_copy = time.perf_counter()
        k_main = ctx.copy_k()
        tris = ctx.copy_triangles()
        t_copy = time.perf_counter() - t_copy
        parity = {"checked": False, "bit_exact": None} if args.no_verify else verify(preset, st, k_main, tris)
        del tris
        lcount = ctx.copy_edge_triangles().astype(np.uint64)
        gs["sum_l2"] = int((lcount * lcount).sum())
        gs["sum_l"] = int(lcount.sum())
        del lcount
        other = None
        if not args.no_other_variant:
            # the other variant on the same graph (same launch shapes), timed the
            # same way; its k must equal the headline variant's
            ov = "ondemand" if wl["variant"] == "direct" else "direct"
            try:
                ctx.reset()
                ctx.decompose(ov, wl["parts"])
                t2 = time.perf_counter()
                for _ in range(args.steps):
                    ctx.reset()
                    so = ctx.decompose(ov, wl["parts"])
                t3 = time.perf_counter()
                ms_o = (t3 - t2) / args.steps * 1e3
                other = {"variant": ov, "ms_per_step": ms_o, "value": so["four_cliques"] / (ms_o / 1e3),
                         "k_equal": bool(np.array_equal(ctx.copy_k(), k_main))}
            except nuc34.Nuc34Error as e:  # e.g. the direct lists do not fit
                other = {"variant": ov, "error": str(e)}
        del k_main
        ctx.close()
        rates = None if args.no_rates else nuc34.device_rates(0)
        batched = None if args.no_batch else run_batched(nuc34, args.batch_graphs, autotune=args.autotune)
        serving = (run_serving(nuc34, args.serve_seconds, args.serve_rate, args.serve_large_share,
                               small_workers=args.serve_small_workers) if args.serve_seconds > 0 else None)
        cpu = None
        if not args.no_cpu_baseline:
            cpu = run_same_input_pair(nuc34, tmp, args.cpu_threads)
            rec = restatement_record()
            if rec and args.workload == "rmat20k30":
                rec["speedup_vs_this_run"] = rec["decomp_s"] / ((t1 - t0) / args.steps)
                cpu["full_workload_restatement"] = rec
        kern = {}
        for s_ in steps:
            for name, k in s_["kernels"].items():
                a = kern.setdefault(name, {"ms": 0.0, "launches": 0})
                a["ms"] += k["ms"]
                a["launches"] += k["launches"]
        dom = max(kern, key=lambda k: kern[k]["ms"])
        launches_per_step = kern[dom]["launches"] / args.steps
        traffic = None if args.no_traffic else run_traffic(args.workload, path, dom, launches_per_step, tmp, tuning)

    ms = (t1 - t0) / args.steps * 1e3
    exact = parity["bit_exact"]
    value = st["four_cliques"] / (ms / 1e3) if exact is not False else None
    end_to_end = {"parse_ms": t_parse * 1e3, "load_ms": t_load * 1e3, "decompose_ms": ms,
                  "copy_k_and_triangles_ms": t_copy * 1e3}
    end_to_end["total_ms"] = sum(end_to_end.values())
    # achieved algorithmic GB/s of every timed kernel group (intersection, support, lists, peel)
    kernel_gbs = {k: round(algorithmic_bytes(k, gs, st) / (v["ms"] / args.steps / 1e3) / 1e9, 1)
                  for k, v in kern.items() if v["ms"] > 0 and algorithmic_bytes(k, gs, st) > 0}
    # physical GB/s (PMC fetch + write bytes of the profiled step over the timed step's kernel time)
    kernel_physical_gbs = ({k: round(b / (kern[k]["ms"] / args.steps / 1e3) / 1e9, 1)
                            for k, b in traffic["group_bytes_per_step"].items() if kern.get(k, {}).get("ms", 0) > 0}
                           if traffic else None)
    per_launch_ms = kern[dom]["ms"] / kern[dom]["launches"]
    bytes_per_launch = algorithmic_bytes(dom, gs, st) / launches_per_step
    achieved = bytes_per_launch / (per_launch_ms / 1e3) / 1e9
    rec = {
        "metric": "4-cliques/s (decomposition wall time, bit-exact k)",
        "value": value,
        "unit": "4-cliques/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded generator, gpu-nucleus34_amd/_build/nuc34_gen)",
        "config": {"workload": wl["desc"], "variant": wl["variant"], "parts": wl["parts"],
                   "n": gs["n"], "m_oriented": gs["m"], "triangles": st["triangles"],
                   "four_cliques": st["four_cliques"], "K_max": st["k_max"], "rounds": st["rounds"],
                   "sum_l2": gs["sum_l2"], "parity": parity, "tuning": tuning,
                   "phase_ms": {k: round(st[k], 3) for k in ("ms_triangles", "ms_support", "ms_k4_list",
                                                           "ms_incidence", "ms_peel", "ms_total")},
                   "kernel_ms_per_step": {k: round(v["ms"] / args.steps, 3) for k, v in kern.items()},
                   "kernel_launches_per_step": {k: v["launches"] / args.steps for k, v in kern.items()},
                   "kernel_gbs": kernel_gbs, "kernel_physical_gbs": kernel_physical_gbs,
                   "end_to_end_ms": {k: round(v, 3) for k, v in end_to_end.items()}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_detail": traffic,
                     "bytes_per_launch": bytes_per_launch, "ms_per_launch": per_launch_ms,
                     "launches_per_step": launches_per_step,
                     "atomics": peel_atomics(st, rates, traffic)},
        "device_rates": rates,
        "other_variant": other,
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(rec))
    if exact is False:
        print(f"results differ from the golden values: {parity['mismatched']}", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
