"""Python host mirror of the (3,4)-nucleus C ABI (include/nuc34.h).

Loads the in-tree build of libnuc34.so (gpu-nucleus34_amd/_build) and exposes
the reference's operator surface: dataset loading with the reference loader
semantics, and a device context that runs orientation, triangles, incidence,
support and the direct or on-demand peel on the GPU. There is no CPU fallback:
if the library is missing, importing this module raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(os.path.dirname(_HERE), "_build")
LIB_PATH = os.path.join(BUILD_DIR, "libnuc34.so")

OK, E_OOM, E_OVERFLOW, E_HIP, E_INVALID, E_IO, E_STATE = range(7)
DIRECT, ONDEMAND, AUTO, AUTO_FIT = 0, 1, 2, 3
VARIANTS = {"direct": DIRECT, "ondemand": ONDEMAND, "auto": AUTO, "auto_fit": AUTO_FIT}


class Nuc34Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ServeConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget", C.c_uint64), ("large_fraction", C.c_double),
                ("large_edges", C.c_uint64), ("max_batch_edges", C.c_uint64), ("max_batch_graphs", C.c_uint32),
                ("variant", C.c_int), ("batch_window_ms", C.c_double), ("large_workers", C.c_uint32),
                ("large_wedges", C.c_uint64), ("small_workers", C.c_uint32)]


class JobInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("variant", C.c_int32), ("n", C.c_uint64), ("m", C.c_uint64),
                ("triangles", C.c_uint64), ("four_cliques", C.c_uint64), ("k_max", C.c_int32),
                ("batch_graphs", C.c_uint32), ("batch_id", C.c_uint64), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_done", C.c_double), ("error", C.c_char * 96)]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["error"] = self.error.decode()
        return d


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("batches", C.c_uint64), ("splits", C.c_uint64), ("busy_s", C.c_double * 2)]


class DeviceRates(C.Structure):
    _fields_ = [("copy_gbps", C.c_double), ("atomic_gops", C.c_double), ("load_l2_gops", C.c_double),
                ("load_hbm_gops", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("peel_debit_grid", C.c_uint32), ("peel_tile_cap", C.c_uint32),
                ("peel_rounds_per_launch", C.c_uint32), ("k4_list_grid_cap", C.c_uint32),
                ("tri_tile_edges", C.c_uint32), ("peel_finish_grid", C.c_uint32),
                ("batch_peel", C.c_uint32), ("graph_peel_max_tris", C.c_uint32),
                ("tri_schedule", C.c_uint32)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("ms", C.c_double), ("launches", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [
        ("n", C.c_uint64), ("m", C.c_uint64), ("triangles", C.c_uint64), ("four_cliques", C.c_uint64),
        ("incidence_entries", C.c_uint64), ("rounds", C.c_uint64), ("unpeeled", C.c_uint64),
        ("peak_bytes", C.c_uint64), ("k_max", C.c_int32), ("variant", C.c_int32),
        ("ms_orient", C.c_double), ("ms_triangles", C.c_double), ("ms_incidence", C.c_double),
        ("ms_support", C.c_double), ("ms_k4_list", C.c_double), ("ms_peel", C.c_double),
        ("ms_total", C.c_double), ("n_kernels", C.c_uint32), ("kernels", KernelStat * 8),
        ("debits", C.c_uint64), ("ms_tri_order", C.c_double), ("ms_peel_init", C.c_double),
    ]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_ if f not in ("kernels", "n_kernels")}
        d["kernels"] = {self.kernels[i].name.decode(): {"ms": self.kernels[i].ms, "launches": self.kernels[i].launches}
                        for i in range(self.n_kernels)}
        return d


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libnuc34.so not built at {LIB_PATH}; run __graft_entry__.build() or make -C gpu-nucleus34_amd")
    lib = C.CDLL(LIB_PATH)
    u32p, i32p, u64p = C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(C.c_uint64)
    vp = C.c_void_p
    sig = {
        "nuc34_graph_load": (C.c_int, [C.c_char_p, C.POINTER(vp), C.c_char_p, C.c_size_t]),
        "nuc34_graph_from_pairs": (C.c_int, [u32p, u32p, C.c_uint64, C.POINTER(vp)]),
        "nuc34_graph_info": (C.c_int, [vp, u32p, u64p, C.POINTER(C.c_int)]),
        "nuc34_graph_copy": (C.c_int, [vp, u32p, u32p]),
        "nuc34_graph_orient_host": (C.c_int, [vp]),
        "nuc34_graph_free": (None, [vp]),
        "nuc34_gen_preset": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
        "nuc34_gen_batch_graph": (C.c_int, [C.c_uint64, C.c_uint64, C.POINTER(vp)]),
        "nuc34_gen_rmat": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                     C.c_uint32, C.c_uint32, C.POINTER(vp)]),
        "nuc34_ctx_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(vp)]),
        "nuc34_ctx_destroy": (None, [vp]),
        "nuc34_last_error": (C.c_char_p, [vp]),
        "nuc34_load_csr": (C.c_int, [vp, u32p, u32p, C.c_uint32, C.c_uint64, C.c_int]),
        "nuc34_graph_dims": (C.c_int, [vp, u32p, u64p]),
        "nuc34_copy_oriented": (C.c_int, [vp, u32p, u32p]),
        "nuc34_triangles": (C.c_int, [vp, C.c_uint32, u64p]),
        "nuc34_copy_triangles": (C.c_int, [vp, C.c_uint64, C.c_uint64, u32p]),
        "nuc34_copy_edge_triangles": (C.c_int, [vp, u32p]),
        "nuc34_support": (C.c_int, [vp, u64p]),
        "nuc34_copy_support": (C.c_int, [vp, i32p]),
        "nuc34_decompose": (C.c_int, [vp, C.c_int, C.c_uint32, C.POINTER(Stats)]),
        "nuc34_copy_k": (C.c_int, [vp, i32p]),
        "nuc34_copy_fourcliques": (C.c_int, [vp, C.c_uint64, C.c_uint64, u32p]),
        "nuc34_get_stats": (C.c_int, [vp, C.POINTER(Stats)]),
        "nuc34_reset": (C.c_int, [vp]),
        "nuc34_load_batch": (C.c_int, [vp, C.c_uint32, u32p, u64p, u32p, u32p, C.c_int]),
        "nuc34_load_batch_ptrs": (C.c_int, [vp, C.c_uint32, u32p, u64p, C.POINTER(u32p), C.POINTER(u32p), C.c_int]),
        "nuc34_batch_graph": (C.c_int, [vp, C.c_uint32, u64p, u64p, u64p, u64p]),
        "nuc34_batch_slices": (C.c_int, [vp, u64p, u64p]),
        "nuc34_batch_rounds": (C.c_int, [vp, u32p]),
        "nuc34_set_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "nuc34_measure_device": (C.c_int, [C.c_int, C.POINTER(DeviceRates)]),
        "nuc34_get_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "nuc34_autotune": (C.c_int, [vp, C.c_int, C.POINTER(Tuning), C.POINTER(C.c_double)]),
        "nuc34_device_info": (C.c_int, [C.c_int, C.c_char_p, C.c_size_t]),
        "nuc34_decompose_noalloc": (C.c_int, [C.c_int, u32p, u32p, C.c_uint32, C.c_uint64, C.c_int, C.c_uint64,
                                              u32p, i32p, C.POINTER(Stats)]),
        "nuc34_serve_create": (C.c_int, [C.POINTER(ServeConfig), C.POINTER(vp)]),
        "nuc34_serve_destroy": (None, [vp]),
        "nuc34_serve_now": (C.c_double, [vp]),
        "nuc34_serve_submit": (C.c_int, [vp, u32p, u32p, C.c_uint32, C.c_uint64, u64p]),
        "nuc34_serve_wait": (C.c_int, [vp, C.c_uint64, C.c_int64, C.POINTER(JobInfo)]),
        "nuc34_serve_take": (C.c_int, [vp, C.c_uint64, u32p, i32p]),
        "nuc34_serve_get_stats": (C.c_int, [vp, C.POINTER(ServeStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()


def lib() -> C.CDLL:
    return _lib


def _u32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _i32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


@dataclass
class HostGraph:
    """A dataset as the reference loads it: undirected CSR (oriented=False) or a
    pre-oriented .grh CSR (oriented=True)."""
    O: np.ndarray
    E: np.ndarray
    oriented: bool

    @property
    def n(self) -> int:
        return len(self.O) - 1

    @property
    def m(self) -> int:
        return len(self.E)


def _take_graph(h) -> HostGraph:
    n, m, o = C.c_uint32(), C.c_uint64(), C.c_int()
    _lib.nuc34_graph_info(h, C.byref(n), C.byref(m), C.byref(o))
    O = np.empty(n.value + 1, dtype=np.uint32)
    E = np.empty(max(m.value, 1), dtype=np.uint32)
    _lib.nuc34_graph_copy(h, _u32(O), _u32(E))
    _lib.nuc34_graph_free(h)
    return HostGraph(O, E[: m.value].copy(), bool(o.value))


def load_dataset(path: str) -> HostGraph:
    """graph_t::from_edges / deserialize by extension (.mtx, .edges, .txt, .grh)."""
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    rc = _lib.nuc34_graph_load(os.fsencode(path), C.byref(h), err, 512)
    if rc != OK:
        raise Nuc34Error(rc, err.value.decode(errors="replace"))
    return _take_graph(h)


def graph_from_pairs(src, dst) -> HostGraph:
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    h = C.c_void_p()
    rc = _lib.nuc34_graph_from_pairs(_u32(src), _u32(dst), len(src), C.byref(h))
    if rc != OK:
        raise Nuc34Error(rc, "graph_from_pairs failed")
    return _take_graph(h)


def gen_preset(name: str) -> HostGraph:
    h = C.c_void_p()
    rc = _lib.nuc34_gen_preset(name.encode(), C.byref(h))
    if rc != OK:
        raise Nuc34Error(rc, f"unknown preset {name}")
    return _take_graph(h)


def gen_batch_graph(seed: int, index: int) -> HostGraph:
    h = C.c_void_p()
    rc = _lib.nuc34_gen_batch_graph(seed, index, C.byref(h))
    if rc != OK:
        raise Nuc34Error(rc, "batch generator failed")
    return _take_graph(h)


def gen_rmat(scale, edge_factor, a, b, c, seed, planted_cliques=0, planted_size=0) -> HostGraph:
    h = C.c_void_p()
    rc = _lib.nuc34_gen_rmat(scale, edge_factor, a, b, c, seed, planted_cliques, planted_size, C.byref(h))
    if rc != OK:
        raise Nuc34Error(rc, "rmat generator failed")
    return _take_graph(h)


def device_rates(device: int = 0) -> dict:
    """Measured copy GB/s and random atomic / load rates (10^9 per second)."""
    r = DeviceRates()
    rc = _lib.nuc34_measure_device(device, C.byref(r))
    if rc != OK:
        raise Nuc34Error(rc, "nuc34_measure_device failed")
    return {f: getattr(r, f) for f, _ in r._fields_}


def device_info(device: int = 0) -> dict:
    buf = C.create_string_buffer(4096)
    rc = _lib.nuc34_device_info(device, buf, 4096)
    if rc != OK:
        raise Nuc34Error(rc, "no HIP device")
    return dict(line.split("=", 1) for line in buf.value.decode().splitlines() if "=" in line)


class Context:
    """A device context: owns every device buffer of one decomposition."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._h = C.c_void_p()
        rc = _lib.nuc34_ctx_create(device, hbm_budget, C.byref(self._h))
        if rc != OK:
            raise Nuc34Error(rc, "nuc34_ctx_create failed: " + (_lib.nuc34_last_error(None) or b"").decode())

    def close(self):
        if self._h:
            _lib.nuc34_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != OK:
            raise Nuc34Error(rc, _lib.nuc34_last_error(self._h).decode(errors="replace"))

    def load_csr(self, O, E, oriented: bool):
        O = np.ascontiguousarray(O, dtype=np.uint32)
        E = np.ascontiguousarray(E, dtype=np.uint32)
        Ep = E if len(E) else np.zeros(1, dtype=np.uint32)
        self._check(_lib.nuc34_load_csr(self._h, _u32(O), _u32(Ep), len(O) - 1, len(E), int(bool(oriented))))

    def load_graph(self, g: HostGraph):
        self.load_csr(g.O, g.E, g.oriented)

    def load_batch(self, graphs, oriented: bool = False):
        """Load many graphs (HostGraph or (O, E) pairs) as one disjoint union."""
        Os, Es, ns, ms = [], [], [], []
        for g in graphs:
            O, E = (g.O, g.E) if isinstance(g, HostGraph) else g
            Os.append(np.ascontiguousarray(O, dtype=np.uint32))
            Es.append(np.ascontiguousarray(E, dtype=np.uint32))
            ns.append(len(O) - 1)
            ms.append(len(E))
        n = np.array(ns, dtype=np.uint32)
        m = np.array(ms, dtype=np.uint64)
        if not ns:
            z = np.zeros(1, np.uint32)
            self._check(_lib.nuc34_load_batch(self._h, 0, _u32(n), m.ctypes.data_as(C.POINTER(C.c_uint64)),
                                              _u32(z), _u32(z), int(bool(oriented))))
        else:
            # one pointer per graph: the library uploads each into place (the
            # arrays in Os/Es stay alive for the call)
            Op = (C.POINTER(C.c_uint32) * len(ns))(*[_u32(a) for a in Os])
            Ep = (C.POINTER(C.c_uint32) * len(ns))(*[_u32(a) if len(a) else None for a in Es])
            self._check(_lib.nuc34_load_batch_ptrs(self._h, len(ns), _u32(n), m.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                   Op, Ep, int(bool(oriented))))
        self._batch = len(ns)

    def batch_graph(self, g: int, with_triangles: bool = True) -> dict:
        vf, ef, tf, tc = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        if with_triangles:
            self._check(_lib.nuc34_batch_graph(self._h, g, C.byref(vf), C.byref(ef), C.byref(tf), C.byref(tc)))
        else:
            self._check(_lib.nuc34_batch_graph(self._h, g, C.byref(vf), C.byref(ef), None, None))
        return {"vertex_first": vf.value, "edge_first": ef.value, "tri_first": tf.value, "tri_count": tc.value}

    def batch_slices(self):
        """(edge_first, tri_first) of every graph of the loaded batch, count + 1 entries each."""
        ef = np.empty(self._batch + 1, dtype=np.uint64)
        tf = np.empty(self._batch + 1, dtype=np.uint64)
        p = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint64))
        self._check(_lib.nuc34_batch_slices(self._h, p(ef), p(tf)))
        return ef, tf

    def batch_rounds(self) -> np.ndarray:
        """Peel rounds of every graph of the batch, when the last decomposition
        peeled it one workgroup per graph (each equals the graph's own rounds)."""
        r = np.empty(max(self._batch, 1), dtype=np.uint32)
        self._check(_lib.nuc34_batch_rounds(self._h, _u32(r)))
        return r[:self._batch]

    def dims(self):
        n, m = C.c_uint32(), C.c_uint64()
        self._check(_lib.nuc34_graph_dims(self._h, C.byref(n), C.byref(m)))
        return n.value, m.value

    def oriented_csr(self):
        n, m = self.dims()
        O = np.empty(n + 1, dtype=np.uint32)
        E = np.empty(max(m, 1), dtype=np.uint32)
        self._check(_lib.nuc34_copy_oriented(self._h, _u32(O), _u32(E)))
        return O, E[:m]

    def triangles(self, parts: int = 1) -> int:
        T = C.c_uint64()
        self._check(_lib.nuc34_triangles(self._h, parts, C.byref(T)))
        return T.value

    def copy_triangles(self, first: int = 0, count: int | None = None) -> np.ndarray:
        T = self.triangles()
        if count is None:
            count = T - first
        out = np.empty((max(count, 1), 3), dtype=np.uint32)
        self._check(_lib.nuc34_copy_triangles(self._h, first, count, _u32(out)))
        return out[:count]

    def copy_edge_triangles(self) -> np.ndarray:
        """Triangles on each oriented edge (m entries, any role)."""
        self.triangles()
        _, m = self.dims()
        out = np.empty(max(m, 1), dtype=np.uint32)
        self._check(_lib.nuc34_copy_edge_triangles(self._h, _u32(out)))
        return out[:m]

    def support(self) -> int:
        Q = C.c_uint64()
        self._check(_lib.nuc34_support(self._h, C.byref(Q)))
        return Q.value

    def copy_support(self) -> np.ndarray:
        T = self.triangles()
        out = np.empty(max(T, 1), dtype=np.int32)
        self._check(_lib.nuc34_copy_support(self._h, _i32(out)))
        return out[:T]

    def decompose(self, variant="auto", parts: int = 1) -> dict:
        v = VARIANTS[variant] if isinstance(variant, str) else int(variant)
        st = Stats()
        self._check(_lib.nuc34_decompose(self._h, v, parts, C.byref(st)))
        return st.as_dict()

    def copy_k(self) -> np.ndarray:
        T = self.triangles()
        out = np.empty(max(T, 1), dtype=np.int32)
        self._check(_lib.nuc34_copy_k(self._h, _i32(out)))
        return out[:T]

    def copy_fourcliques(self, first: int = 0, count: int | None = None) -> np.ndarray:
        if count is None:
            count = self.stats()["four_cliques"] - first
        out = np.empty((max(count, 1), 4), dtype=np.uint32)
        self._check(_lib.nuc34_copy_fourcliques(self._h, first, count, _u32(out)))
        return out[:count]

    def tuning(self) -> dict:
        t = Tuning()
        self._check(_lib.nuc34_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **fields):
        t = Tuning()
        _lib.nuc34_get_tuning(self._h, C.byref(t))
        for k, v in fields.items():
            setattr(t, k, v)
        self._check(_lib.nuc34_set_tuning(self._h, C.byref(t)))

    def autotune(self, variant="auto"):
        """Times candidate launch shapes on the loaded graph and keeps the
        fastest; returns (tuning dict, decomposition ms)."""
        t, ms = Tuning(), C.c_double()
        v = VARIANTS[variant] if isinstance(variant, str) else int(variant)
        self._check(_lib.nuc34_autotune(self._h, v, C.byref(t), C.byref(ms)))
        return t.as_dict(), ms.value

    def reset(self):
        """Drop every stage after the loaded graph (buffers go to the reuse cache)."""
        self._check(_lib.nuc34_reset(self._h))

    def stats(self) -> dict:
        st = Stats()
        self._check(_lib.nuc34_get_stats(self._h, C.byref(st)))
        return st.as_dict()


def decompose_graph(g: HostGraph, variant="auto", parts: int = 1, device: int = 0, hbm_budget: int = 0):
    """One-shot: returns (triangles (T,3), k (T,), stats)."""
    with Context(device, hbm_budget) as ctx:
        ctx.load_graph(g)
        st = ctx.decompose(variant, parts)
        return ctx.copy_triangles(), ctx.copy_k(), st


def decompose_noalloc(g, variant="auto", tri_capacity: int | None = None, device: int = 0):
    """nuc34_decompose_noalloc: one undirected graph decomposed with every device
    buffer in the library's per-device scratch array (no device allocation).
    Returns (triangles (T,3), k (T,), stats); tri_capacity defaults to the
    graph's wedge count, an upper bound on T."""
    O, E = (g.O, g.E) if isinstance(g, HostGraph) else g
    O = np.ascontiguousarray(O, dtype=np.uint32)
    E = np.ascontiguousarray(E, dtype=np.uint32)
    if tri_capacity is None:
        d = np.diff(O.astype(np.int64))
        tri_capacity = int((d * (d - 1) // 2).sum()) // 3 + 1
    abc = np.empty((max(tri_capacity, 1), 3), dtype=np.uint32)
    k = np.empty(max(tri_capacity, 1), dtype=np.int32)
    st = Stats()
    v = VARIANTS[variant] if isinstance(variant, str) else int(variant)
    rc = _lib.nuc34_decompose_noalloc(device, _u32(O), _u32(E if len(E) else O), len(O) - 1, len(E), v,
                                      tri_capacity, _u32(abc), _i32(k), C.byref(st))
    if rc != OK:
        err = Nuc34Error(rc, (_lib.nuc34_last_error(None) or b"").decode(errors="replace"))
        err.stats = st.as_dict()
        raise err
    T = st.triangles
    return abc[:T], k[:T], st.as_dict()


class Server:
    """Decomposition service (nuc34_serve_*): submit undirected graphs from any
    thread, wait for and take their results. Small jobs are batched as disjoint
    unions by `small_workers` workers (a job arriving while a batch runs starts
    on the next free one), large ones run alone on `large_workers` workers; the
    workers share the device through their own contexts and streams."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, large_fraction: float = 0.0, large_edges: int = 0,
                 max_batch_edges: int = 0, max_batch_graphs: int = 0, variant="auto", batch_window_ms: float = 0.0,
                 large_workers: int = 0, large_wedges: int = 0, small_workers: int = 0):
        cfg = ServeConfig(device, hbm_budget, large_fraction, large_edges, max_batch_edges, max_batch_graphs,
                          VARIANTS[variant] if isinstance(variant, str) else int(variant), batch_window_ms,
                          large_workers, large_wedges, small_workers)
        self._h = C.c_void_p()
        rc = _lib.nuc34_serve_create(C.byref(cfg), C.byref(self._h))
        if rc != OK:
            raise Nuc34Error(rc, "nuc34_serve_create failed")
        self._T = {}

    def close(self):
        if self._h:
            _lib.nuc34_serve_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self) -> float:
        return _lib.nuc34_serve_now(self._h)

    def submit(self, g) -> int:
        O, E = (g.O, g.E) if isinstance(g, HostGraph) else g
        O = np.ascontiguousarray(O, dtype=np.uint32)
        E = np.ascontiguousarray(E, dtype=np.uint32)
        job = C.c_uint64()
        rc = _lib.nuc34_serve_submit(self._h, _u32(O), _u32(E if len(E) else O), len(O) - 1, len(E), C.byref(job))
        if rc != OK:
            raise Nuc34Error(rc, "submit rejected")
        return job.value

    def wait(self, job: int, timeout_ms: int = -1) -> dict | None:
        info = JobInfo()
        rc = _lib.nuc34_serve_wait(self._h, job, timeout_ms, C.byref(info))
        if rc == E_STATE:
            return None
        if rc != OK:
            raise Nuc34Error(rc, "wait failed")
        return info.as_dict()

    def take(self, job: int, triangles: int):
        """Triangles (T,3) with the graph's own oriented edge ids, and k (T,); forgets the job."""
        abc = np.empty((max(triangles, 1), 3), dtype=np.uint32)
        k = np.empty(max(triangles, 1), dtype=np.int32)
        rc = _lib.nuc34_serve_take(self._h, job, _u32(abc), _i32(k))
        if rc != OK:
            raise Nuc34Error(rc, "take failed (job unknown or not done)")
        return abc[:triangles], k[:triangles]

    def stats(self) -> dict:
        st = ServeStats()
        _lib.nuc34_serve_get_stats(self._h, C.byref(st))
        return {"submitted": st.submitted, "completed": st.completed, "failed": st.failed, "batches": st.batches,
                "splits": st.splits, "busy_s": [st.busy_s[0], st.busy_s[1]]}
