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


class MatchStats(ctypes.Structure):
    _fields_ = [("matched_vertices", ctypes.c_int64), ("augmentations", ctypes.c_int64),
                ("blossom_edges", ctypes.c_int64), ("edges_scanned", ctypes.c_int64),
                ("blossoms", ctypes.c_int64), ("blossom_vertices", ctypes.c_int64),
                ("phases", ctypes.c_int32), ("levels", ctypes.c_int32),
                ("aug_path_edges", ctypes.c_int64), ("expand_edges", ctypes.c_int64),
                ("blossom_scan_edges", ctypes.c_int64), ("reused_tree_vertices", ctypes.c_int64),
                ("reset_tree_vertices", ctypes.c_int64)]


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
        for fn in ("oracle_cpu_msbfs_dir", "oracle_cpu_msbfs_bitpar"):
            getattr(lib, fn).argtypes = [I32, _P32, _P32, _P32, _P32, _P32, I32, _P32, _P32, I32]
            getattr(lib, fn).restype = I64
        lib.oracle_edge_map.argtypes = [_P32, I64, _P32, _P32]
        lib.oracle_edge_map.restype = I64
        lib.oracle_exclusive_scan.argtypes = [_P32, I64, _P64]
        lib.oracle_exclusive_scan.restype = I64
        lib.oracle_match_forest.argtypes = [I32, _P32, _P32, _P32, ctypes.POINTER(MatchStats), I32, I32]
        lib.oracle_match_forest.restype = I64
        lib.oracle_match_edmonds.argtypes = [I32, _P32, _P32, _P32]
        lib.oracle_match_edmonds.restype = I64
        lib.oracle_match_hk.argtypes = [I32, _P32, _P32, _P32]
        lib.oracle_match_hk.restype = I64
        lib.oracle_match_validate.argtypes = [I32, _P32, _P32, _P32]
        lib.oracle_match_validate.restype = I64
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


def transpose(row_offsets, col_indices):
    """In-edge CSR of a directed CSR: (in_offsets, in_cols), each in-list in
    ascending source order."""
    ro = np.asarray(row_offsets, dtype=np.int64)
    ci = np.asarray(col_indices, dtype=np.int64)
    n = ro.size - 1
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(ro))
    order = np.lexsort((src, ci))
    in_off = np.zeros(n + 1, dtype=np.int64)
    np.add.at(in_off, ci + 1, 1)
    return np.cumsum(in_off).astype(np.int32), src[order].astype(np.int32)


def msbfs_cpu(row_offsets, col_indices, sources, threads: int, mode: str = "bitpar", transposed=None):
    """Multi-threaded CPU restatements (the timed CPU baseline): "dir" = one
    direction-optimizing BFS per source, threads over sources; "bitpar" = 64
    sources per word, each level parallel over vertices in chunks of 64.
    transposed: (in_offsets, in_cols) of a directed graph (None: symmetric).
    Returns (levels, parents, traversed_edges), as msbfs()."""
    lib = load()
    ro = np.ascontiguousarray(row_offsets, dtype=np.int32)
    ci = np.ascontiguousarray(col_indices, dtype=np.int32)
    if ci.size == 0:
        ci = np.zeros(1, dtype=np.int32)
    io, ic = (ro, ci) if transposed is None else (np.ascontiguousarray(transposed[0], dtype=np.int32),
                                                   np.ascontiguousarray(transposed[1], dtype=np.int32))
    if ic.size == 0:
        ic = np.zeros(1, dtype=np.int32)
    src = np.ascontiguousarray(sources, dtype=np.int32)
    n, S = ro.size - 1, src.size
    lv = np.empty((S, n), dtype=np.int32)
    pa = np.empty((S, n), dtype=np.int32)
    if S == 0:
        return lv, pa, 0
    fn = {"dir": lib.oracle_cpu_msbfs_dir, "bitpar": lib.oracle_cpu_msbfs_bitpar}[mode]
    e = fn(n, _p(ro), _p(ci), _p(io), _p(ic), _p(src), S, _p(lv), _p(pa), max(1, threads))
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


def _csr(row_offsets, col_indices):
    ro = np.ascontiguousarray(row_offsets, dtype=np.int32)
    ci = np.ascontiguousarray(col_indices, dtype=np.int32)
    if ci.size == 0:
        ci = np.zeros(1, dtype=np.int32)
    return ro, ci, ro.size - 1


def match_forest(row_offsets, col_indices, init_mate=None, max_phases: int = 0, reuse: bool = True):
    """Deterministic alternating-forest matching: (mate, MatchStats); -1 = exposed.

    reuse keeps the trees of still-exposed roots across phases (the default,
    as in the engine); reuse=False regrows the forest every phase.
    max_phases > 0 stops after that many phases (a bounded timing sample)."""
    lib = load()
    ro, ci, n = _csr(row_offsets, col_indices)
    mate = np.full(max(n, 1), -1, dtype=np.int32)
    if init_mate is not None:
        mate[:n] = np.asarray(init_mate, dtype=np.int32)
    st = MatchStats()
    lib.oracle_match_forest(n, _p(ro), _p(ci), _p(mate), ctypes.byref(st), max_phases, 1 if reuse else 0)
    return mate[:n], st


def match_edmonds(row_offsets, col_indices):
    """Maximum cardinality matching (Edmonds' blossom algorithm), -1 = exposed."""
    lib = load()
    ro, ci, n = _csr(row_offsets, col_indices)
    mate = np.full(max(n, 1), -1, dtype=np.int32)
    lib.oracle_match_edmonds(n, _p(ro), _p(ci), _p(mate))
    return mate[:n]


def match_hk(row_offsets, col_indices):
    """Maximum matching of a bipartite graph (Hopcroft-Karp); None if not bipartite."""
    lib = load()
    ro, ci, n = _csr(row_offsets, col_indices)
    mate = np.full(max(n, 1), -1, dtype=np.int32)
    r = lib.oracle_match_hk(n, _p(ro), _p(ci), _p(mate))
    return None if r < 0 else mate[:n]


def match_violations(row_offsets, col_indices, mate):
    lib = load()
    ro, ci, n = _csr(row_offsets, col_indices)
    m = np.ascontiguousarray(mate, dtype=np.int32)
    return int(lib.oracle_match_validate(n, _p(ro), _p(ci), _p(m)))
