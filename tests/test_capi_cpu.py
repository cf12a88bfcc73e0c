"""CPU-side checks of the C ABI library (no GPU calls that compute).

* lib/libxbmsbfs.so loads and exports every function include/xb/xb_capi.h declares;
* CSR validation errors are reported with status codes before any device work;
* without a visible GPU, a valid graph is refused with XB_ERR_NO_DEVICE (the
  product never falls back to a CPU path);
* the text loader rejects malformed files with XB_ERR_IO.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import xbmsbfs as xb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "xb", h) for h in os.listdir(os.path.join(ROOT, "include", "xb"))
           if h.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        names |= set(re.findall(r"^\s*(?:int32_t|void|uint64_t|int64_t)\s+(xb_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_header_declarations_match_binding():
    assert set(declared_functions()) == set(xb.EXPORTED_SYMBOLS)


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    lib = xb.load_engine()
    assert hasattr(lib, name), f"{name} not exported"


def test_opts_default():
    o = xb.default_opts()
    assert o.words_per_tile in (1, 2, 4, 8) and o.alpha > 0
    assert o.pull_cut_small <= o.pull_cut_medium <= o.pull_cut_hub


def _create(ro, ci, directed=False):
    lib = xb.load_engine()
    h = ctypes.c_void_p()
    ro = np.ascontiguousarray(ro, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    rc = lib.xb_graph_create_csr(xb._ptr(ro), ro.size, xb._ptr(ci), ci.size, 1 if directed else 0, ctypes.byref(h))
    buf = ctypes.create_string_buffer(512)
    lib.xb_last_error(buf, 512)
    return rc, buf.value.decode()


def _has_gpu():
    # Probe without initialising any HIP runtime in the test process.
    return os.path.exists("/dev/kfd")


@pytest.mark.parametrize("ro,ci,msg", [
    ([0, 2, 1], [1, 0], "monotone"),
    ([1, 2, 3], [1, 0, 0], "row_offsets[0]"),
    ([0, 1, 2], [1, 0, 0], "number of column"),
    ([0, 1, 2], [1, 5], "out of range"),
    ([0, 1, 2], [1, -1], "out of range"),
])
def test_bad_csr_rejected(ro, ci, msg):
    rc, err = _create(ro, ci)
    assert rc == xb.XB_ERR_BAD_CSR and msg in err


def test_null_and_empty_arguments():
    lib = xb.load_engine()
    h = ctypes.c_void_p()
    assert lib.xb_graph_create_csr(None, 0, None, 0, 0, ctypes.byref(h)) == xb.XB_ERR_INVALID_ARG
    ro = np.array([0, 0], dtype=np.int32)
    assert lib.xb_graph_create_csr(xb._ptr(ro), 2, None, 0, 7, ctypes.byref(h)) == xb.XB_ERR_INVALID_ARG
    assert lib.xb_msbfs_run(None, None, 1, None, None, None, None) == xb.XB_ERR_INVALID_ARG


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device refusal")
def test_valid_graph_without_gpu_is_refused():
    rc, err = _create([0, 1, 2], [1, 0])
    assert rc == xb.XB_ERR_NO_DEVICE, err


def test_text_loader_errors(tmp_path):
    lib = xb.load_engine()
    h = ctypes.c_void_p()
    missing = str(tmp_path / "nope.txt").encode()
    assert lib.xb_graph_load_text(missing, missing, 0, ctypes.byref(h)) == xb.XB_ERR_IO
    bad = tmp_path / "bad.txt"
    bad.write_text("0 1 x 2\n")
    good = tmp_path / "col.txt"
    good.write_text("1 0\n")
    assert lib.xb_graph_load_text(str(bad).encode(), str(good).encode(), 0, ctypes.byref(h)) == xb.XB_ERR_IO
    ro = tmp_path / "ro.txt"
    ro.write_text("0 1\n3\n")  # row_offsets[n] != number of column indices
    assert lib.xb_graph_load_text(str(ro).encode(), str(good).encode(), 0, ctypes.byref(h)) == xb.XB_ERR_BAD_CSR


def test_server_defaults_and_null_arguments():
    lib = xb.load_engine()
    so = xb.ServerOpts()
    assert lib.xb_server_opts_default(ctypes.byref(so)) == xb.XB_OK
    assert so.slots >= 1 and so.max_queued >= 1 and so.flags & xb.XB_SERVE_COUNT_EDGES
    h = ctypes.c_void_p()
    assert lib.xb_server_create(None, ctypes.byref(so), None, ctypes.byref(h)) == xb.XB_ERR_INVALID_ARG
    rid = ctypes.c_uint64()
    src = np.zeros(1, dtype=np.int32)
    assert lib.xb_server_submit(None, xb._ptr(src), 1, ctypes.byref(rid)) == xb.XB_ERR_INVALID_ARG
    assert lib.xb_server_wait(None, 1, 0, None) == xb.XB_ERR_INVALID_ARG
    assert lib.xb_server_release(None, 1) == xb.XB_ERR_INVALID_ARG
    assert lib.xb_server_stats(None, None) == xb.XB_ERR_INVALID_ARG
    assert lib.xb_server_destroy(None) == xb.XB_OK


def test_match_null_arguments():
    lib = xb.load_engine()
    mate = np.zeros(4, dtype=np.int32)
    st = xb.MatchStats()
    assert lib.xb_match_run(None, None, xb._ptr(mate), 0, ctypes.byref(st)) == xb.XB_ERR_INVALID_ARG


def test_binary_csr_round_trip_and_corruption(tmp_path):
    g = xb.gen_rmat(10, 8, 0.57, 0.19, 0.19, seed=3)
    path = str(tmp_path / "g.xbcsr")
    xb.save_csr_binary(path, g.row_offsets, g.col_indices, directed=True)
    ro, ci, directed = xb.load_csr_binary(path)
    assert directed and np.array_equal(ro, g.row_offsets) and np.array_equal(ci, g.col_indices)
    assert os.path.getsize(path) == 40 + 4 * (g.row_offsets.size + g.col_indices.size)
    raw = bytearray(open(path, "rb").read())
    # A flipped column bit breaks the checksum.
    bad = bytearray(raw)
    bad[-1] ^= 1
    (tmp_path / "bad.xbcsr").write_bytes(bytes(bad))
    with pytest.raises(xb.XbError) as ei:
        xb.load_csr_binary(str(tmp_path / "bad.xbcsr"))
    assert ei.value.code == xb.XB_ERR_IO and "checksum" in str(ei.value)
    # Truncated arrays.
    (tmp_path / "short.xbcsr").write_bytes(bytes(raw[:-8]))
    with pytest.raises(xb.XbError) as ei:
        xb.load_csr_binary(str(tmp_path / "short.xbcsr"))
    assert ei.value.code == xb.XB_ERR_IO
    # Wrong magic.
    (tmp_path / "magic.xbcsr").write_bytes(b"NOTACSR\0" + bytes(raw[8:]))
    with pytest.raises(xb.XbError) as ei:
        xb.load_csr_binary(str(tmp_path / "magic.xbcsr"))
    assert ei.value.code == xb.XB_ERR_IO
    # Structure is validated where the graph is built: a non-monotone file loads
    # as arrays but is refused by xb_graph_load_binary before any device work.
    xb.save_csr_binary(str(tmp_path / "nm.xbcsr"), [0, 2, 1], [1, 0], directed=False)
    lib = xb.load_engine()
    h = ctypes.c_void_p()
    assert lib.xb_graph_load_binary(str(tmp_path / "nm.xbcsr").encode(), ctypes.byref(h)) == xb.XB_ERR_BAD_CSR


def test_text_loader_parses_large_files_in_parallel(tmp_path):
    # Above the 4 MiB threshold the parser splits the file across threads; a
    # malformed token late in the file must still be reported, and a valid
    # file must get as far as the device check (no device here).
    g = xb.gen_rmat(15, 16, 0.57, 0.19, 0.19, seed=4)
    rp, cp = tmp_path / "ro.txt", tmp_path / "ci.txt"
    np.savetxt(rp, g.row_offsets, fmt="%d")
    np.savetxt(cp, g.col_indices, fmt="%d")
    assert os.path.getsize(cp) > (4 << 20)
    lib = xb.load_engine()
    h = ctypes.c_void_p()
    rc = lib.xb_graph_load_text(str(rp).encode(), str(cp).encode(), 0, ctypes.byref(h))
    if not _has_gpu():
        assert rc == xb.XB_ERR_NO_DEVICE
    txt = cp.read_text()
    cut = len(txt) - 1000
    cut = txt.index("\n", cut)
    (tmp_path / "bad.txt").write_text(txt[:cut] + "\n12x3" + txt[cut:])
    rc = lib.xb_graph_load_text(str(rp).encode(), str(tmp_path / "bad.txt").encode(), 0, ctypes.byref(h))
    assert rc == xb.XB_ERR_IO


def test_noalloc_null_arguments():
    """The allocation-free entry point validates before touching a device."""
    lib = xb.load_engine()
    src = np.zeros(1, dtype=np.int32)
    st = xb.RunStats()
    assert lib.xb_msbfs_run_noalloc(None, xb._ptr(src), 1, None, None, None, ctypes.byref(st)) == xb.XB_ERR_INVALID_ARG
