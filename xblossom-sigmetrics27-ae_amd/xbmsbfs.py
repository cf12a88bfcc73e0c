"""ctypes binding of the MI355X multi-source BFS engine (lib/libxbmsbfs.so).

Mirrors the reference's C++ entry points for the search path:

* ``HGraph(row_offsets_path, col_indices_path)`` --
  ``zblossom::HGraph`` (apps/xb/src/include/graph/hgraph_gpu.h:37-48), here
  also constructible from in-memory int32 CSR arrays;
* ``XBlossomGpuEngine10(graph, path_table_buffer_ratio)`` --
  ``zblossom::XBlossomGpuEngine10(const HGraph&, size_t)``
  (apps/xb/src/include/xbsm_gpu/xblossom_gpu_engine_10.h:25), gaining
  ``MultiSourceBFS(sources, levels, parents)``.

Every call goes through the C ABI in include/xb/xb_capi.h.  There is no CPU
fallback: if the shared library is missing or no GPU is visible, calls raise.
The generator functions wrap lib/libxbgen.so (host-only).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
ENGINE_LIB = os.path.join(LIB_DIR, "libxbmsbfs.so")
GEN_LIB = os.path.join(LIB_DIR, "libxbgen.so")

XB_OK = 0
XB_ERR_INVALID_ARG = -1
XB_ERR_BAD_CSR = -2
XB_ERR_IO = -3
XB_ERR_NOMEM = -4
XB_ERR_DEVICE = -5
XB_ERR_SOURCE_RANGE = -6
XB_ERR_BUDGET = -7
XB_ERR_UNSUPPORTED = -8
XB_ERR_NO_DEVICE = -9
XB_ERR_TIMEOUT = -10

XB_GRAPH_SYMMETRIC = 0
XB_GRAPH_DIRECTED = 1

XB_MODE_AUTO = 0
XB_MODE_PLAIN = 1
XB_MODE_BATCHED = 2

XB_OUT_DEVICE = 1
XB_PROFILE = 2
XB_NO_LEVELS = 4
XB_FIXED_TEAMS = 8
XB_MATCH_REGROW = 16
XB_NO_PIPELINE = 32
XB_BATCH_EMIT_INSIDE = 64

# Every symbol include/xb/xb_capi.h declares.
EXPORTED_SYMBOLS = (
    "xb_opts_default",
    "xb_graph_create_csr",
    "xb_graph_load_text",
    "xb_csr_save",
    "xb_csr_load_header",
    "xb_csr_load",
    "xb_graph_load_binary",
    "xb_graph_save_binary",
    "xb_graph_info",
    "xb_graph_destroy",
    "xb_msbfs_run",
    "xb_msbfs_run_noalloc",
    "xb_count_traversed",
    "xb_device_checksum",
    "xb_test_edge_map",
    "xb_batch_create",
    "xb_batch_run",
    "xb_batch_destroy",
    "xb_msbfs_run_batched",
    "xb_device_alloc",
    "xb_device_free",
    "xb_memcpy_d2h",
    "xb_memcpy_h2d",
    "xb_last_error",
    "xb_autotune",
    "xb_server_opts_default",
    "xb_server_create",
    "xb_server_submit",
    "xb_server_wait",
    "xb_server_release",
    "xb_server_stats",
    "xb_server_destroy",
    "xb_match_run",
)

XB_SERVE_COUNT_EDGES = 1
# Bytes of the static device array behind xb_msbfs_run_noalloc (per GPU).
XB_NOALLOC_SCRATCH_BYTES = 64 << 20


class XbError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"xb error {code}: {msg}")
        self.code = code
        self.msg = msg


class Opts(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("words_per_tile", ctypes.c_int32),
        ("alpha", ctypes.c_int32),
        ("pull_cut_small", ctypes.c_int32),
        ("pull_cut_medium", ctypes.c_int32),
        ("pull_cut_hub", ctypes.c_int32),
        ("hub_chunk", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 8),
    ]


# Kernel classes of xb_run_stats (XB_K_* in include/xb/xb_capi.h).
KERNEL_CLASSES = ("memset", "seed", "push", "commit", "pull", "pull_hub", "clear", "emit_heavy", "emit",
                  "plain_init", "plain_push", "persistent")
K_CLASSES = 17
# Kernel classes of xb_match_stats (XB_K_MT_*).
MATCH_KERNEL_CLASSES = ("",) * 12 + ("mt_scan", "mt_expand", "mt_tree", "mt_blossom", "mt_init")


class RunStats(ctypes.Structure):
    _fields_ = [
        ("ms_total", ctypes.c_double),
        ("kernel_ms", ctypes.c_double * K_CLASSES),
        ("kernel_bytes", ctypes.c_double * K_CLASSES),
        ("kernel_launches", ctypes.c_int64 * K_CLASSES),
        ("max_level", ctypes.c_int32),
        ("tiles", ctypes.c_int32),
        ("push_levels", ctypes.c_int32),
        ("pull_levels", ctypes.c_int32),
        ("fallback_sources", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 11),
    ]

    def kernels(self) -> dict:
        """Per kernel class: summed HIP-event ms, launches and algorithmic bytes."""
        out = {}
        for i, name in enumerate(KERNEL_CLASSES):
            if self.kernel_launches[i]:
                out[name] = {"ms": self.kernel_ms[i], "launches": int(self.kernel_launches[i]),
                             "bytes": self.kernel_bytes[i]}
        return out

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_
             if k not in ("reserved", "kernel_ms", "kernel_bytes", "kernel_launches")}
        d["kernels"] = self.kernels()
        return d


TUNABLES = ("words_per_tile", "alpha", "pull_cut_small", "pull_cut_medium", "pull_cut_hub", "hub_chunk")


class TuneTrial(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in TUNABLES] + [("ms_per_source", ctypes.c_double)]


class TuneReport(ctypes.Structure):
    _fields_ = [("trials", ctypes.c_int32), ("passes", ctypes.c_int32),
                ("default_ms_per_source", ctypes.c_double), ("best_ms_per_source", ctypes.c_double),
                ("trial", TuneTrial * 128)]

    def as_dict(self) -> dict:
        return {"passes": self.passes, "default_ms_per_source": self.default_ms_per_source,
                "best_ms_per_source": self.best_ms_per_source,
                "trials": [dict({k: getattr(t, k) for k in TUNABLES}, ms_per_source=t.ms_per_source)
                           for t in self.trial[: self.trials]]}


class ServerOpts(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int32), ("max_queued", ctypes.c_int32), ("budget_bytes", ctypes.c_int64),
                ("flags", ctypes.c_uint32), ("reserved", ctypes.c_int32 * 7)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("S", ctypes.c_int32), ("n", ctypes.c_int64),
                ("levels", ctypes.c_void_p), ("parents", ctypes.c_void_p), ("edges", ctypes.c_int64),
                ("t_submit_us", ctypes.c_double), ("t_admit_us", ctypes.c_double),
                ("t_start_us", ctypes.c_double), ("t_done_us", ctypes.c_double),
                ("tiles", ctypes.c_int32), ("reserved", ctypes.c_int32 * 7)]


class ServerCounters(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("failed", ctypes.c_int64),
                ("released", ctypes.c_int64), ("tiles", ctypes.c_int64), ("tile_rows", ctypes.c_int64),
                ("budget_bytes", ctypes.c_int64), ("used_bytes", ctypes.c_int64),
                ("peak_used_bytes", ctypes.c_int64), ("queued", ctypes.c_int32),
                ("running_tiles", ctypes.c_int32), ("reserved", ctypes.c_int32 * 8)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class MatchStats(ctypes.Structure):
    _fields_ = [("matched_vertices", ctypes.c_int64), ("augmentations", ctypes.c_int64),
                ("blossom_edges", ctypes.c_int64), ("edges_scanned", ctypes.c_int64),
                ("blossoms", ctypes.c_int64), ("blossom_vertices", ctypes.c_int64),
                ("phases", ctypes.c_int32), ("levels", ctypes.c_int32), ("ms_total", ctypes.c_double),
                ("kernel_ms", ctypes.c_double * K_CLASSES), ("kernel_bytes", ctypes.c_double * K_CLASSES),
                ("kernel_launches", ctypes.c_int64 * K_CLASSES),
                ("aug_path_edges", ctypes.c_int64), ("expand_edges", ctypes.c_int64),
                ("blossom_scan_edges", ctypes.c_int64), ("reused_tree_vertices", ctypes.c_int64),
                ("reset_tree_vertices", ctypes.c_int64), ("reserved", ctypes.c_int32 * 8)]

    def kernels(self) -> dict:
        return {name: {"ms": self.kernel_ms[i], "bytes": self.kernel_bytes[i],
                       "launches": int(self.kernel_launches[i])}
                for i, name in enumerate(MATCH_KERNEL_CLASSES) if name and self.kernel_launches[i]}

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_
                if k not in ("reserved", "kernel_ms", "kernel_bytes", "kernel_launches")}


_lib = None
_gen = None

_P32 = ctypes.POINTER(ctypes.c_int32)


def _ptr(a):
    """int32 pointer to a numpy array, or a raw device address (int)."""
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.cast(ctypes.c_void_p(a), _P32)
    return a.ctypes.data_as(_P32)


def load_engine():
    """Loads lib/libxbmsbfs.so; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ENGINE_LIB):
        raise XbError(XB_ERR_UNSUPPORTED, f"{ENGINE_LIB} not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(ENGINE_LIB)
    I32, U32, I64, VP = ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64, ctypes.c_void_p
    lib.xb_opts_default.argtypes = [ctypes.POINTER(Opts)]
    lib.xb_graph_create_csr.argtypes = [_P32, I64, _P32, I64, U32, ctypes.POINTER(VP)]
    lib.xb_graph_load_text.argtypes = [ctypes.c_char_p, ctypes.c_char_p, U32, ctypes.POINTER(VP)]
    lib.xb_graph_info.argtypes = [VP, ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(U32)]
    lib.xb_csr_save.argtypes = [ctypes.c_char_p, _P32, I64, _P32, I64, U32]
    lib.xb_csr_load_header.argtypes = [ctypes.c_char_p, ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(U32)]
    lib.xb_csr_load.argtypes = [ctypes.c_char_p, _P32, I64, _P32, I64]
    lib.xb_graph_load_binary.argtypes = [ctypes.c_char_p, ctypes.POINTER(VP)]
    lib.xb_graph_save_binary.argtypes = [VP, ctypes.c_char_p]
    lib.xb_graph_destroy.argtypes = [VP]
    lib.xb_msbfs_run.argtypes = [VP, _P32, I32, _P32, _P32, ctypes.POINTER(Opts), ctypes.POINTER(RunStats)]
    lib.xb_msbfs_run_noalloc.argtypes = [VP, _P32, I32, _P32, _P32, ctypes.POINTER(Opts),
                                         ctypes.POINTER(RunStats)]
    lib.xb_count_traversed.argtypes = [VP, _P32, I32, ctypes.POINTER(I64)]
    lib.xb_device_checksum.argtypes = [_P32, I64, ctypes.POINTER(ctypes.c_uint64)]
    lib.xb_test_edge_map.argtypes = [_P32, I32, _P32, _P32, ctypes.POINTER(I64)]
    PP32 = ctypes.POINTER(_P32)
    lib.xb_batch_create.argtypes = [ctypes.POINTER(VP), I32, PP32, _P32, PP32, PP32, ctypes.POINTER(Opts),
                                    ctypes.POINTER(VP)]
    lib.xb_batch_run.argtypes = [VP, ctypes.POINTER(RunStats)]
    lib.xb_batch_destroy.argtypes = [VP]
    lib.xb_msbfs_run_batched.argtypes = [ctypes.POINTER(VP), I32, PP32, _P32, PP32, PP32, ctypes.POINTER(Opts),
                                         ctypes.POINTER(RunStats)]
    lib.xb_device_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(VP)]
    lib.xb_device_free.argtypes = [VP]
    lib.xb_memcpy_d2h.argtypes = [VP, VP, ctypes.c_size_t]
    lib.xb_memcpy_h2d.argtypes = [VP, VP, ctypes.c_size_t]
    lib.xb_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.xb_autotune.argtypes = [VP, _P32, I32, ctypes.POINTER(Opts), ctypes.POINTER(Opts),
                                ctypes.POINTER(TuneReport)]
    lib.xb_server_opts_default.argtypes = [ctypes.POINTER(ServerOpts)]
    lib.xb_server_create.argtypes = [VP, ctypes.POINTER(ServerOpts), ctypes.POINTER(Opts), ctypes.POINTER(VP)]
    lib.xb_server_submit.argtypes = [VP, _P32, I32, ctypes.POINTER(ctypes.c_uint64)]
    lib.xb_server_wait.argtypes = [VP, ctypes.c_uint64, I32, ctypes.POINTER(RequestInfo)]
    lib.xb_server_release.argtypes = [VP, ctypes.c_uint64]
    lib.xb_server_stats.argtypes = [VP, ctypes.POINTER(ServerCounters)]
    lib.xb_server_destroy.argtypes = [VP]
    lib.xb_match_run.argtypes = [VP, _P32, _P32, U32, ctypes.POINTER(MatchStats)]
    for name in EXPORTED_SYMBOLS:
        getattr(lib, name).restype = ctypes.c_int32
    _lib = lib
    return lib


def _check(rc: int):
    if rc != XB_OK:
        buf = ctypes.create_string_buffer(1024)
        _lib.xb_last_error(buf, 1024)
        raise XbError(rc, buf.value.decode(errors="replace"))


def default_opts(**kw) -> Opts:
    lib = load_engine()
    o = Opts()
    _check(lib.xb_opts_default(ctypes.byref(o)))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class HGraph:
    """Device-resident CSR graph (int32 rowOffsets / colIndices)."""

    def __init__(self, row_offsets=None, col_indices=None, directed: bool = False):
        lib = load_engine()
        self._h = ctypes.c_void_p()
        flags = XB_GRAPH_DIRECTED if directed else XB_GRAPH_SYMMETRIC
        if isinstance(row_offsets, (str, os.PathLike)):
            _check(lib.xb_graph_load_text(os.fsencode(row_offsets), os.fsencode(col_indices), flags,
                                          ctypes.byref(self._h)))
        else:
            ro = np.ascontiguousarray(row_offsets, dtype=np.int32)
            ci = np.ascontiguousarray(col_indices if col_indices is not None else [], dtype=np.int32)
            _check(lib.xb_graph_create_csr(_ptr(ro), ro.size, _ptr(ci), ci.size, flags, ctypes.byref(self._h)))
        n, m, f = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_uint32()
        _check(lib.xb_graph_info(self._h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(f)))
        self.nnodes, self.nedges, self.directed = n.value, m.value, bool(f.value & 1)

    @classmethod
    def from_binary(cls, path: str) -> "HGraph":
        """Loads a binary CSR file (xb_graph_load_binary)."""
        g = cls.__new__(cls)
        lib = load_engine()
        g._h = ctypes.c_void_p()
        _check(lib.xb_graph_load_binary(os.fsencode(path), ctypes.byref(g._h)))
        n, m, f = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_uint32()
        _check(lib.xb_graph_info(g._h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(f)))
        g.nnodes, g.nedges, g.directed = n.value, m.value, bool(f.value & 1)
        return g

    def save_binary(self, path: str):
        _check(load_engine().xb_graph_save_binary(self._h, os.fsencode(path)))

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _check(_lib.xb_graph_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class XBlossomGpuEngine10:
    """Engine bound to one graph; ``MultiSourceBFS`` is the batched search."""

    def __init__(self, graph: HGraph, path_table_buffer_ratio: int = 80):
        self.graph = graph
        self.path_table_buffer_ratio = path_table_buffer_ratio
        self.last_stats = RunStats()
        self.tunables = {}  # set by Autotune; explicit keyword tunables override them

    def Autotune(self, sample, **base) -> TuneReport:
        """Tunes words per tile, alpha and the pull cut-offs on one tile of
        ``sample`` sources; later MultiSourceBFS calls use the result."""
        best, rep = autotune(self.graph, sample, **base)
        self.tunables = best
        return rep

    def MultiSourceBFS(self, sources, levels=None, parents=None, *, mode: int = XB_MODE_AUTO,
                       device_out: bool = False, profile: bool = False, want_levels: bool = True,
                       pipeline: bool = True, **tunables):
        """Runs len(sources) BFS traversals.

        Host mode (default): returns (levels, parents) as S x n int32 numpy
        arrays.  Device mode: ``levels``/``parents`` are device addresses
        (ints, e.g. ``tensor.data_ptr()``) of S x n int32 buffers; a run of
        several tiles then alternates two workspaces and streams unless
        ``pipeline`` is False (XB_NO_PIPELINE).
        """
        lib = load_engine()
        src = np.ascontiguousarray(sources, dtype=np.int32)
        S, n = int(src.size), self.graph.nnodes
        flags = (XB_OUT_DEVICE if device_out else 0) | (XB_PROFILE if profile else 0)
        if not want_levels:
            flags |= XB_NO_LEVELS
        if not pipeline:
            flags |= XB_NO_PIPELINE
        if not device_out:
            if parents is None:
                parents = np.empty((S, n), dtype=np.int32)
            if levels is None and want_levels:
                levels = np.empty((S, n), dtype=np.int32)
        o = default_opts(mode=mode, flags=flags, **{**self.tunables, **tunables})
        st = RunStats()
        _check(lib.xb_msbfs_run(self.graph.handle, _ptr(src), S, _ptr(levels), _ptr(parents), ctypes.byref(o),
                                ctypes.byref(st)))
        self.last_stats = st
        return levels, parents

    def CountTraversed(self, levels_dev: int, S: int) -> int:
        lib = load_engine()
        out = ctypes.c_int64()
        _check(lib.xb_count_traversed(self.graph.handle, _ptr(levels_dev), S, ctypes.byref(out)))
        return out.value

    def FindMaximumMatch(self, initial_matching=None) -> np.ndarray:
        """FindMaximumMatch of apps/xb/src/include/xbsm_gpu/max_match_engine.h:
        84-113: entry v is v's mate, nnodes for an exposed vertex (the
        reference GPU engine's sentinel).  The result is a maximum matching,
        blossoms included (``last_match_stats.blossoms`` counts them)."""
        n = self.graph.nnodes
        init = None
        if initial_matching is not None:
            init = np.array(initial_matching, dtype=np.int32).reshape(-1)
            if init.size != n:
                raise XbError(XB_ERR_INVALID_ARG, "initial matching must have nnodes entries")
            init[init == n] = -1
        mate, st = match(self.graph, init, profile=True)
        self.last_match_stats = st
        # Counters and times add up over the engine's calls, as the
        # reference's members do (xblossom_gpu_engine_10.h:99-116).
        acc = self.__dict__.setdefault("_match_totals", dict.fromkeys(
            ("aug", "expand", "blossom_edges", "blossoms", "reused", "reset", "t_aug", "t_expand", "t_blossom",
             "t_pre"), 0))
        acc["aug"] += st.aug_path_edges
        acc["expand"] += st.expand_edges
        acc["blossom_edges"] += st.blossom_scan_edges
        acc["blossoms"] += st.blossoms
        acc["reused"] += st.reused_tree_vertices
        acc["reset"] += st.reset_tree_vertices
        acc["t_aug"] += st.kernel_ms[12] + st.kernel_ms[14]
        acc["t_expand"] += st.kernel_ms[13]
        acc["t_blossom"] += st.kernel_ms[15]
        acc["t_pre"] += st.kernel_ms[16]
        mate[mate < 0] = n
        return mate

    # Statistics and timing of apps/xb/src/include/xbsm_gpu/xblossom_gpu_engine_10.h:39-95.
    def _tot(self, k):
        return self.__dict__.get("_match_totals", {}).get(k, 0)

    def GetTimingSummary(self):
        """(augmenting_path, expand, blossom, total) device time in SECONDS,
        as the reference returns them (nanoseconds / 1e9 under _ms names)."""
        a, e, b = self._tot("t_aug") / 1e3, self._tot("t_expand") / 1e3, self._tot("t_blossom") / 1e3
        return a, e, b, a + e + b

    def PrintTimingSummary(self):
        a, e, b, t = self.GetTimingSummary()
        print(f"XBlossomGpuEngine10(MI355X: Reuse+ParBlossom+MinKey) Timing Summary:\nAugmenting Path: {a}\n"
              f"Expand: {e}\nBlossom: {b}\nPreprocessing Time: {self._tot('t_pre') / 1e3}\nTotal Time: {t}")

    def NumEdgesProcessed(self) -> int:
        return self._tot("aug") + self._tot("expand") + self._tot("blossom_edges")

    def NumAugmentingPathEdgesProcessed(self) -> int:
        return self._tot("aug")

    def NumExpandEdgesProcessed(self) -> int:
        return self._tot("expand")

    def NumBlossomEdgesProcessed(self) -> int:
        return self._tot("blossom_edges")

    def NumBlossoms(self) -> int:
        return self._tot("blossoms")

    def NumReusedTreeNodes(self) -> int:
        return self._tot("reused")

    def NumResetTreeNodes(self) -> int:
        return self._tot("reset")


class DeviceBuffer:
    """Device int32 array from the library's own allocator (no PyTorch needed).

    In a process that also uses PyTorch-ROCm, import torch before the first
    call into this module: both then share torch's HIP runtime."""

    def __init__(self, shape, fill=None):
        lib = load_engine()
        self.shape = tuple(int(x) for x in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.count = int(np.prod(self.shape)) if self.shape else 1
        self._p = ctypes.c_void_p()
        _check(lib.xb_device_alloc(max(self.count, 1) * 4, ctypes.byref(self._p)))
        if fill is not None:
            self.upload(np.full(self.count, fill, dtype=np.int32))

    @property
    def ptr(self) -> int:
        return self._p.value

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.int32).reshape(-1)
        assert a.size == self.count
        _check(load_engine().xb_memcpy_h2d(self._p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))

    def rows(self, r0: int, r1: int) -> np.ndarray:
        """Copies rows [r0, r1) of a 2-D buffer to the host."""
        n = self.shape[1]
        out = np.empty((r1 - r0, n), dtype=np.int32)
        src = ctypes.c_void_p(self._p.value + r0 * n * 4)
        _check(load_engine().xb_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), src, out.nbytes))
        return out

    def numpy(self) -> np.ndarray:
        out = np.empty(self.count, dtype=np.int32)
        _check(load_engine().xb_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self._p, out.nbytes))
        return out.reshape(self.shape)

    def checksum(self) -> int:
        out = ctypes.c_uint64()
        _check(load_engine().xb_device_checksum(_ptr(self.ptr), self.count, ctypes.byref(out)))
        return out.value

    def free(self):
        if self._p:
            load_engine().xb_device_free(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr_array(items):
    arr = (_P32 * len(items))()
    for i, a in enumerate(items):
        arr[i] = _ptr(a) if a is not None else None
    return arr


class Batch:
    """A planned multi-graph batch: one persistent, hipGraph-captured launch.

    graphs: list of HGraph; sources: list of int32 arrays; levels/parents: lists
    of device addresses (ints) of S_i x n_i int32 tables (levels may be None with
    want_levels=False).  Each run re-splits the workgroups between the graphs'
    teams from the run's team clocks unless fixed_teams is set.  A graph whose
    sources fit one tile is emitted by standalone launches captured after the
    persistent kernel, unless emit_inside is set (XB_BATCH_EMIT_INSIDE)."""

    def __init__(self, graphs, sources, levels, parents, want_levels=True, profile=False, fixed_teams=False,
                 emit_inside=False, **tunables):
        lib = load_engine()
        self._keep = [np.ascontiguousarray(s, dtype=np.int32) for s in sources]
        G = len(graphs)
        self.graphs = graphs
        gh = (ctypes.c_void_p * G)(*[g.handle.value for g in graphs])
        S = np.array([s.size for s in self._keep], dtype=np.int32)
        self._S = S
        flags = (XB_OUT_DEVICE | (0 if want_levels else XB_NO_LEVELS) | (XB_PROFILE if profile else 0)
                 | (XB_FIXED_TEAMS if fixed_teams else 0) | (XB_BATCH_EMIT_INSIDE if emit_inside else 0))
        o = default_opts(flags=flags, **tunables)
        self._h = ctypes.c_void_p()
        lv = _ptr_array(levels) if want_levels else None
        _check(lib.xb_batch_create(gh, G, _ptr_array(self._keep), _ptr(S), lv, _ptr_array(parents),
                                   ctypes.byref(o), ctypes.byref(self._h)))
        self.last_stats = RunStats()

    def run(self) -> RunStats:
        st = RunStats()
        _check(load_engine().xb_batch_run(self._h, ctypes.byref(st)))
        self.last_stats = st
        return st

    def close(self):
        if self._h:
            _check(load_engine().xb_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching query server over one device-resident graph.

    submit(sources) -> request id; wait(id) -> RequestInfo with device
    pointers to the request's S x n level and parent tables (valid until
    release(id)).  Rows of several requests share bit-parallel tiles, tiles
    on different slots run concurrently, and a request is admitted once its
    tables fit in the device budget."""

    def __init__(self, graph: HGraph, slots: int = 4, budget_bytes: int = 0, count_edges: bool = True,
                 max_queued: int = 0, **tunables):
        lib = load_engine()
        so = ServerOpts()
        _check(lib.xb_server_opts_default(ctypes.byref(so)))
        so.slots = slots
        so.budget_bytes = budget_bytes
        so.flags = XB_SERVE_COUNT_EDGES if count_edges else 0
        if max_queued:
            so.max_queued = max_queued
        eo = default_opts(**tunables)
        self.graph = graph
        self._h = ctypes.c_void_p()
        _check(lib.xb_server_create(graph.handle, ctypes.byref(so), ctypes.byref(eo), ctypes.byref(self._h)))

    def submit(self, sources) -> int:
        src = np.ascontiguousarray(sources, dtype=np.int32)
        rid = ctypes.c_uint64()
        _check(load_engine().xb_server_submit(self._h, _ptr(src), src.size, ctypes.byref(rid)))
        return rid.value

    def wait(self, rid: int, timeout_ms: int = -1) -> RequestInfo:
        info = RequestInfo()
        _check(load_engine().xb_server_wait(self._h, rid, timeout_ms, ctypes.byref(info)))
        return info

    def release(self, rid: int):
        _check(load_engine().xb_server_release(self._h, rid))

    def counters(self) -> ServerCounters:
        c = ServerCounters()
        _check(load_engine().xb_server_stats(self._h, ctypes.byref(c)))
        return c

    def close(self):
        if self._h:
            _check(load_engine().xb_server_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(graph: HGraph, sample, **base):
    """Runs xb_autotune; returns (tuned tunables as a dict, TuneReport)."""
    src = np.ascontiguousarray(sample, dtype=np.int32)
    o = default_opts(**base)
    best = Opts()
    rep = TuneReport()
    _check(load_engine().xb_autotune(graph.handle, _ptr(src), src.size, ctypes.byref(o), ctypes.byref(best),
                                     ctypes.byref(rep)))
    return {k: int(getattr(best, k)) for k in TUNABLES}, rep


def match(graph: HGraph, init_mate=None, mate_out=None, *, device: bool = False, profile: bool = False,
          regrow: bool = False):
    """xb_match_run: (mate, MatchStats), -1 for exposed vertices.

    Host mode: ``init_mate`` is None or n int32 entries and a new numpy mate
    array is returned.  Device mode: ``init_mate`` (or None) and ``mate_out``
    are device addresses of n int32 entries.  ``regrow`` resets the whole
    forest every phase (XB_MATCH_REGROW) instead of keeping the trees of
    still-exposed roots."""
    lib = load_engine()
    n = graph.nnodes
    st = MatchStats()
    prof = (XB_PROFILE if profile else 0) | (XB_MATCH_REGROW if regrow else 0)
    if device:
        if mate_out is None:
            raise XbError(XB_ERR_INVALID_ARG, "device mode needs mate_out")
        _check(lib.xb_match_run(graph.handle, _ptr(init_mate), _ptr(mate_out), XB_OUT_DEVICE | prof,
                                ctypes.byref(st)))
        return mate_out, st
    init = None if init_mate is None else np.ascontiguousarray(init_mate, dtype=np.int32)
    if init is not None and init.size != n:
        raise XbError(XB_ERR_INVALID_ARG, "init_mate must have n entries")
    out = np.empty(max(n, 1), dtype=np.int32)
    _check(lib.xb_match_run(graph.handle, _ptr(init), _ptr(out), prof, ctypes.byref(st)))
    return out[:n], st


def msbfs_run_noalloc(graph: HGraph, sources, levels_dev: int, parents_dev: int, *, mode: int = XB_MODE_AUTO,
                      **tunables) -> RunStats:
    """xb_msbfs_run_noalloc: the S x n tables into caller-owned device buffers
    (addresses as ints), with no allocation inside the call."""
    src = np.ascontiguousarray(sources, dtype=np.int32)
    o = default_opts(mode=mode, flags=XB_OUT_DEVICE, **tunables)
    st = RunStats()
    _check(load_engine().xb_msbfs_run_noalloc(graph.handle, _ptr(src), src.size, _ptr(levels_dev),
                                              _ptr(parents_dev), ctypes.byref(o), ctypes.byref(st)))
    return st


def device_checksum(ptr: int, count: int) -> int:
    """Order-independent checksum of count int32 values at a device address."""
    out = ctypes.c_uint64()
    _check(load_engine().xb_device_checksum(_ptr(int(ptr)), count, ctypes.byref(out)))
    return out.value


def run_batched(graphs, sources, **tunables):
    """Host-output convenience: returns [(levels, parents)] per graph."""
    lib = load_engine()
    srcs = [np.ascontiguousarray(s, dtype=np.int32) for s in sources]
    G = len(graphs)
    gh = (ctypes.c_void_p * G)(*[g.handle.value for g in graphs])
    S = np.array([s.size for s in srcs], dtype=np.int32)
    lv = [np.empty((s.size, g.nnodes), dtype=np.int32) for s, g in zip(srcs, graphs)]
    pa = [np.empty((s.size, g.nnodes), dtype=np.int32) for s, g in zip(srcs, graphs)]
    o = default_opts(**tunables)
    st = RunStats()
    _check(lib.xb_msbfs_run_batched(gh, G, _ptr_array(srcs), _ptr(S), _ptr_array(lv), _ptr_array(pa),
                                    ctypes.byref(o), ctypes.byref(st)))
    return list(zip(lv, pa)), st


def test_edge_map(degrees):
    """Device compaction + edge mapping on a degree list: (owner, offset) per tasklet."""
    lib = load_engine()
    d = np.ascontiguousarray(degrees, dtype=np.int32)
    total = int(d.astype(np.int64).sum())
    owner = np.empty(max(total, 1), dtype=np.int32)
    offset = np.empty(max(total, 1), dtype=np.int32)
    t = ctypes.c_int64()
    _check(lib.xb_test_edge_map(_ptr(d), d.size, _ptr(owner), _ptr(offset), ctypes.byref(t)))
    return owner[:total], offset[:total]


# ---------------------------------------------------------------------------
# Synthetic inputs (host-only generator library)
# ---------------------------------------------------------------------------
class _Csr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("m", ctypes.c_int64), ("rowoff", _P32), ("col", _P32)]


def load_gen():
    global _gen
    if _gen is not None:
        return _gen
    if not os.path.exists(GEN_LIB):
        raise XbError(XB_ERR_UNSUPPORTED, f"{GEN_LIB} not built")
    g = ctypes.CDLL(GEN_LIB)
    D, I32, I64, U64 = ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
    g.xbgen_rmat.argtypes = [I32, I64, D, D, D, U64, I32, I32, ctypes.POINTER(_Csr)]
    g.xbgen_rmat_capped.argtypes = [I32, I32, I64, D, D, D, U64, I32, I32, ctypes.POINTER(_Csr)]
    g.xbgen_chung_lu.argtypes = [I32, D, D, D, D, U64, I32, I32, ctypes.POINTER(_Csr)]
    g.xbgen_free.argtypes = [ctypes.POINTER(_Csr)]
    g.xbgen_sample_sources.argtypes = [I32, _P32, I32, U64, I32, _P32]
    g.xbgen_sample_sources.restype = I32
    g.xbgen_checksum.argtypes = [I32, _P32, I64, _P32]
    g.xbgen_checksum.restype = U64
    _gen = g
    return g


@dataclass
class Csr:
    row_offsets: np.ndarray
    col_indices: np.ndarray

    @property
    def n(self) -> int:
        return self.row_offsets.size - 1

    @property
    def m(self) -> int:
        return self.col_indices.size

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_offsets)

    def checksum(self) -> int:
        g = load_gen()
        return int(g.xbgen_checksum(self.n, _ptr(self.row_offsets), self.m, _ptr(self.col_indices)))


def _take(c: _Csr) -> Csr:
    g = load_gen()
    ro = np.ctypeslib.as_array(c.rowoff, (c.n + 1,)).copy()
    co = np.ctypeslib.as_array(c.col, (max(c.m, 1),))[: c.m].copy()
    g.xbgen_free(ctypes.byref(c))
    return Csr(ro, co)


def save_csr_binary(path: str, row_offsets, col_indices, directed: bool = False):
    """Writes a binary CSR file (host only)."""
    ro = np.ascontiguousarray(row_offsets, dtype=np.int32)
    ci = np.ascontiguousarray(col_indices, dtype=np.int32)
    _check(load_engine().xb_csr_save(os.fsencode(path), _ptr(ro), ro.size, _ptr(ci), ci.size,
                                     XB_GRAPH_DIRECTED if directed else XB_GRAPH_SYMMETRIC))


def load_csr_binary(path: str):
    """Reads a binary CSR file (host only): returns (row_offsets, col_indices, directed)."""
    lib = load_engine()
    np1, m, fl = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_uint32()
    _check(lib.xb_csr_load_header(os.fsencode(path), ctypes.byref(np1), ctypes.byref(m), ctypes.byref(fl)))
    ro = np.empty(np1.value, dtype=np.int32)
    ci = np.empty(max(m.value, 1), dtype=np.int32)
    _check(lib.xb_csr_load(os.fsencode(path), _ptr(ro), np1.value, _ptr(ci), m.value))
    return ro, ci[: m.value], bool(fl.value & XB_GRAPH_DIRECTED)


def gen_rmat(scale: int, edge_factor: int, a: float, b: float, c: float, seed: int, symmetric: bool = True,
             threads: int = 0) -> Csr:
    """R-MAT with 2**scale vertices and edge_factor * 2**scale draws."""
    g = load_gen()
    out = _Csr()
    rc = g.xbgen_rmat(scale, edge_factor * (1 << scale), a, b, c, seed, 1 if symmetric else 0, threads,
                      ctypes.byref(out))
    if rc != 0:
        raise XbError(XB_ERR_INVALID_ARG, f"xbgen_rmat failed ({rc})")
    return _take(out)


def gen_rmat_capped(scale: int, n: int, draws: int, a: float, b: float, c: float, seed: int,
                    symmetric: bool = False, threads: int = 0) -> Csr:
    """R-MAT over 2**scale ids keeping the draws with both endpoints below n."""
    g = load_gen()
    out = _Csr()
    rc = g.xbgen_rmat_capped(scale, n, draws, a, b, c, seed, 1 if symmetric else 0, threads, ctypes.byref(out))
    if rc != 0:
        raise XbError(XB_ERR_INVALID_ARG, f"xbgen_rmat_capped failed ({rc})")
    return _take(out)


def gen_chung_lu(n: int, gamma: float, mean_deg: float, max_deg: float, seed: int, overdraw: float = 1.0,
                 permute: bool = True, threads: int = 0) -> Csr:
    g = load_gen()
    out = _Csr()
    rc = g.xbgen_chung_lu(n, gamma, mean_deg, max_deg, overdraw, seed, 1 if permute else 0, threads,
                          ctypes.byref(out))
    if rc != 0:
        raise XbError(XB_ERR_INVALID_ARG, f"xbgen_chung_lu failed ({rc})")
    return _take(out)


def sample_sources(csr: Csr, S: int, seed: int, degree_weighted: bool = False) -> np.ndarray:
    g = load_gen()
    out = np.empty(S, dtype=np.int32)
    got = g.xbgen_sample_sources(csr.n, _ptr(csr.row_offsets), S, seed, 1 if degree_weighted else 0, _ptr(out))
    return out[:got]


# Synthetic workloads, one per named configuration.  Sizes are the full ones.
def workload_rmat_s16():
    """R-MAT scale 16, edge factor 16, a=.57 b=.19 c=.19, seed 1; 64 uniform sources (seed 2)."""
    g = gen_rmat(16, 16, 0.57, 0.19, 0.19, seed=1)
    return g, sample_sources(g, 64, seed=2)


def workload_amazon_like():
    """Chung-Lu n=403k, gamma 2.3, mean degree 8.4, max ~2,750, seed 3; 256 uniform sources."""
    g = gen_chung_lu(403_000, 2.3, 8.4, 2_800, seed=3, overdraw=1.005)
    return g, sample_sources(g, 256, seed=3)


def workload_uneven_batch(count: int = 32):
    """32 R-MAT graphs: scale uniform in [12, 22] (log-uniform n), edge factor uniform in
    [8, 32] (seed 5); each gets 1-512 sources, log-uniform (seed 6), uniform over
    non-isolated vertices."""
    rg = np.random.default_rng(5)
    rs = np.random.default_rng(6)
    out = []
    for i in range(count):
        scale = int(rg.integers(12, 23))
        ef = int(rg.integers(8, 33))
        S = int(np.floor(np.exp(rs.uniform(0.0, np.log(513.0)))))
        S = max(1, min(512, S))
        g = gen_rmat(scale, ef, 0.57, 0.19, 0.19, seed=1000 + i)
        out.append((g, sample_sources(g, S, seed=2000 + i)))
    return out


def workload_livejournal_like():
    """Chung-Lu n=4.85M, gamma 2.1, mean degree ~14.2, max ~20k, seed 4; 1,024 degree-weighted sources."""
    g = gen_chung_lu(4_850_000, 2.1, 14.2, 22_500, seed=4, overdraw=1.015)
    return g, sample_sources(g, 1024, seed=4, degree_weighted=True)
