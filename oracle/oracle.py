"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never uses
it.  The algorithm and its reference citations are in msbfs_oracle.c.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_P32 = ctypes.POINTER(ctypes.c_int32)
_P64 = ctypes.POINTER(ctypes.c_int64)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        lib = ctypes.CDLL(LIB)
        I32, I64 = ctypes.c_int32, ctypes.c_int64
        lib.oracle_bfs_one.argtypes = [I32, _P32, _P32, I32, _P32, _P32, _P32]
        lib.oracle_bfs_one.restype = I64
        lib.oracle_msbfs.argtypes = [I32, _P32, _P32, _P32, I32, _P32, _P32, I32]
        lib.oracle_msbfs.restype = I64
        lib.oracle_msbfs_bitparallel.argtypes = [I32, _P32, _P32, _P32, I32, _P32, _P32]
        lib.oracle_msbfs_bitparallel.restype = I64
        lib.oracle_edge_map.argtypes = [_P32, I64, _P32, _P32]
        lib.oracle_edge_map.restype = I64
        lib.oracle_exclusive_scan.argtypes = [_P32, I64, _P64]
        lib.oracle_exclusive_scan.restype = I64
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(_P32)


def msbfs(row_offsets, col_indices, sources, threads: int = 1, bitparallel: bool = False):
    """Returns (levels, parents, traversed_edges) for S sources (S x n int32)."""
    lib = load()
    ro = np.ascontiguousarray(row_offsets, dtype=np.int32)
    ci = np.ascontiguousarray(col_indices, dtype=np.int32)
    if ci.size == 0:
        ci = np.zeros(1, dtype=np.int32)
    src = np.ascontiguousarray(sources, dtype=np.int32)
    n, S = ro.size - 1, src.size
    lv = np.empty((S, n), dtype=np.int32)
    pa = np.empty((S, n), dtype=np.int32)
    if S == 0:
        return lv, pa, 0
    if bitparallel:
        e = lib.oracle_msbfs_bitparallel(n, _p(ro), _p(ci), _p(src), S, _p(lv), _p(pa))
    else:
        e = lib.oracle_msbfs(n, _p(ro), _p(ci), _p(src), S, _p(lv), _p(pa), threads)
    return lv, pa, int(e)


def edge_map(degrees):
    """(owner, offset) of every tasklet under the exclusive-scan mapping."""
    lib = load()
    d = np.ascontiguousarray(degrees, dtype=np.int32)
    total = int(d.astype(np.int64).sum())
    owner = np.empty(max(total, 1), dtype=np.int32)
    offset = np.empty(max(total, 1), dtype=np.int32)
    lib.oracle_edge_map(_p(d), d.size, _p(owner), _p(offset))
    return owner[:total], offset[:total]


def exclusive_scan(degrees):
    lib = load()
    d = np.ascontiguousarray(degrees, dtype=np.int32)
    out = np.empty(d.size + 1, dtype=np.int64)
    lib.oracle_exclusive_scan(_p(d), d.size, out.ctypes.data_as(_P64))
    return out
