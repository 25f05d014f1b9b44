"""ORACLE — test infrastructure only (ctypes front end of liboracle34.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module. See oracle34.hpp for the reference code each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle34.so")
CLI_PATH = os.path.join(HERE, "_build", "oracle34")
REF_CLI = os.path.join(HERE, "_ref", "ref34")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u32p, i32p = C.POINTER(C.c_uint32), C.POINTER(C.c_int32)
        vp = C.c_void_p
        for name, res, args in [
            ("o34_load", vp, [C.c_char_p, C.c_int, C.c_char_p, C.c_int]),
            ("o34_from_csr", vp, [u32p, u32p, C.c_uint32, C.c_uint64, C.c_int]),
            ("o34_n", C.c_uint32, [vp]), ("o34_m", C.c_uint64, [vp]),
            ("o34_copy_csr", None, [vp, u32p, u32p]),
            ("o34_run", C.c_int, [vp, C.c_int, C.c_int, C.c_char_p, C.c_int]),
            ("o34_T", C.c_uint64, [vp]), ("o34_Q", C.c_uint64, [vp]), ("o34_kmax", C.c_int32, [vp]),
            ("o34_rounds", C.c_uint64, [vp]),
            ("o34_copy_tris", None, [vp, u32p]), ("o34_copy_k", None, [vp, i32p]),
            ("o34_copy_support", None, [vp, i32p]), ("o34_time", C.c_double, [vp, C.c_int]),
            ("o34_free", None, [vp]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class OracleGraph:
    def __init__(self, handle):
        self._h = handle

    @classmethod
    def load(cls, path: str, orient: bool = True) -> "OracleGraph":
        err = C.create_string_buffer(512)
        h = lib().o34_load(os.fsencode(path), int(orient), err, 512)
        if not h:
            raise RuntimeError(err.value.decode(errors="replace"))
        return cls(h)

    @classmethod
    def from_csr(cls, O, E, orient: bool) -> "OracleGraph":
        O = np.ascontiguousarray(O, dtype=np.uint32)
        E = np.ascontiguousarray(E, dtype=np.uint32)
        Ep = E if len(E) else np.zeros(1, dtype=np.uint32)
        h = lib().o34_from_csr(O.ctypes.data_as(C.POINTER(C.c_uint32)), Ep.ctypes.data_as(C.POINTER(C.c_uint32)),
                               len(O) - 1, len(E), int(orient))
        return cls(h)

    def csr(self):
        L = lib()
        n, m = L.o34_n(self._h), L.o34_m(self._h)
        O = np.empty(n + 1, dtype=np.uint32)
        E = np.empty(max(m, 1), dtype=np.uint32)
        L.o34_copy_csr(self._h, O.ctypes.data_as(C.POINTER(C.c_uint32)), E.ctypes.data_as(C.POINTER(C.c_uint32)))
        return O, E[:m]

    def run(self, mode: str = "host", parts: int = 0) -> dict:
        L = lib()
        err = C.create_string_buffer(512)
        if L.o34_run(self._h, 0 if mode == "host" else 1, parts, err, 512) != 0:
            raise RuntimeError(err.value.decode(errors="replace"))
        T = L.o34_T(self._h)
        tris = np.empty((max(T, 1), 3), dtype=np.uint32)
        k = np.empty(max(T, 1), dtype=np.int32)
        sup = np.empty(max(T, 1), dtype=np.int32)
        L.o34_copy_tris(self._h, tris.ctypes.data_as(C.POINTER(C.c_uint32)))
        L.o34_copy_k(self._h, k.ctypes.data_as(C.POINTER(C.c_int32)))
        L.o34_copy_support(self._h, sup.ctypes.data_as(C.POINTER(C.c_int32)))
        Q = L.o34_Q(self._h)
        return {
            "T": T, "Q": Q, "K_max": L.o34_kmax(self._h), "rounds": L.o34_rounds(self._h),
            "tris": tris[:T], "k": k[:T] if (T and (Q or mode == "ondemand")) else np.zeros(T, np.int32),
            "support": sup[:T],
            "times": [L.o34_time(self._h, i) for i in range(7)],
        }

    def __del__(self):
        if getattr(self, "_h", None):
            lib().o34_free(self._h)
            self._h = None
