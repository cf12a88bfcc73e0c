"""GPU parity of the two-workspace tile pipeline of xb_msbfs_run.

With device outputs, a batched run of several tiles alternates the graph's
workspace and a second one on its own stream, so one tile's level loop
overlaps the previous tile's emit.  Bar: the tables are bit-identical to the
one-stream run (XB_NO_PIPELINE) and to the CPU oracle, including tiles that
fall back to the plain path and tiles deep enough to need level-code planes
4-7, on either workspace, and back-to-back runs into the same buffers.
"""
import numpy as np
import pytest

import graphs as tg
import oracle
import xbmsbfs as xb

pytestmark = pytest.mark.gpu


def _run_device(hg, src, n, pipeline, **kw):
    S = int(src.size)
    lv = xb.DeviceBuffer((S, n), fill=-7)
    pa = xb.DeviceBuffer((S, n), fill=-7)
    eng = xb.XBlossomGpuEngine10(hg)
    eng.MultiSourceBFS(src, lv.ptr, pa.ptr, mode=xb.XB_MODE_BATCHED, device_out=True, pipeline=pipeline, **kw)
    out = (lv.numpy(), pa.numpy(), eng.last_stats)
    lv.free()
    pa.free()
    return out


def _path_graph(n):
    ro = np.zeros(n + 1, dtype=np.int32)
    cols = []
    for v in range(n):
        nb = [u for u in (v - 1, v + 1) if 0 <= u < n]
        cols.extend(nb)
        ro[v + 1] = ro[v] + len(nb)
    return xb.Csr(ro, np.array(cols, dtype=np.int32))


@pytest.fixture(scope="module")
def rmat13():
    g = xb.gen_rmat(13, 8, 0.57, 0.19, 0.19, seed=17)
    return g, xb.HGraph(g.row_offsets, g.col_indices)


@pytest.mark.parametrize("T", [1, 2, 8])
def test_pipelined_tiles_equal_one_stream_and_oracle(rmat13, T):
    g, hg = rmat13
    src = xb.sample_sources(g, 64 * T * 3 + 17, seed=21)  # 4 tiles, the last one partial
    lv_ref, pa_ref, _ = oracle.msbfs(g.row_offsets, g.col_indices, src, threads=8)
    lv_p, pa_p, st = _run_device(hg, src, g.n, True, words_per_tile=T)
    assert st.tiles == 4
    lv_s, pa_s, _ = _run_device(hg, src, g.n, False, words_per_tile=T)
    assert np.array_equal(lv_p, lv_s) and np.array_equal(pa_p, pa_s)
    assert np.array_equal(lv_p, lv_ref), f"levels differ in {(lv_p != lv_ref).sum()} entries"
    assert np.array_equal(pa_p, pa_ref), f"parents differ in {(pa_p != pa_ref).sum()} entries"


def test_pipelined_back_to_back_runs_same_buffers(rmat13):
    """Two pipelined runs into the same buffers: the second run's tiles on the
    second stream start only after the first run's work is done."""
    g, hg = rmat13
    S, n = 3 * 64 + 5, g.n
    src_a = xb.sample_sources(g, S, seed=22)
    src_b = xb.sample_sources(g, S, seed=23)
    lv = xb.DeviceBuffer((S, n), fill=-7)
    pa = xb.DeviceBuffer((S, n), fill=-7)
    eng = xb.XBlossomGpuEngine10(hg)
    for src in (src_a, src_b):
        eng.MultiSourceBFS(src, lv.ptr, pa.ptr, mode=xb.XB_MODE_BATCHED, device_out=True, words_per_tile=1)
    lv_ref, pa_ref, _ = oracle.msbfs(g.row_offsets, g.col_indices, src_b, threads=8)
    assert np.array_equal(lv.numpy(), lv_ref) and np.array_equal(pa.numpy(), pa_ref)


def test_pipelined_plain_fallback_on_both_workspaces():
    """A 600-vertex path beside a star: the two path tiles (every path source
    is deeper than the 8-bit level code) re-run on the plain path, one on
    each workspace; the two star tiles stay bit-parallel."""
    p = _path_graph(600)
    star_ro, star_ci = tg.star(41)
    n1 = p.n
    ro = np.concatenate([p.row_offsets, star_ro[1:] + p.row_offsets[-1]]).astype(np.int32)
    ci = np.concatenate([p.col_indices, star_ci + n1]).astype(np.int32)
    hg = xb.HGraph(ro, ci)
    rng = np.random.default_rng(5)
    src = np.concatenate([rng.integers(0, 600, size=64), rng.integers(0, 600, size=64),
                          rng.integers(n1, n1 + 41, size=64), rng.integers(n1, n1 + 41, size=30)]).astype(np.int32)
    lv_ref, pa_ref, _ = oracle.msbfs(ro, ci, src, threads=8)
    lv, pa, st = _run_device(hg, src, ro.size - 1, True, words_per_tile=1)
    assert st.tiles == 2 and st.fallback_sources == 128
    assert np.array_equal(lv, lv_ref) and np.array_equal(pa, pa_ref)


def test_pipelined_deep_and_shallow_tiles():
    """Level codes 16 and above use planes 4-7, allocated and cleared per
    workspace: a 120-vertex path beside a star, deep and shallow tiles
    alternating over the two workspaces."""
    ro, ci = tg.two_components(120, 1)
    star_ro, star_ci = tg.star(41)
    n1 = ro.size - 1
    ro = np.concatenate([ro, star_ro[1:] + ro[-1]]).astype(np.int32)
    ci = np.concatenate([ci, star_ci + n1]).astype(np.int32)
    hg = xb.HGraph(ro, ci)
    rng = np.random.default_rng(32)
    shallow = rng.integers(n1, n1 + 41, size=128).astype(np.int32)
    deep = rng.integers(0, 120, size=128).astype(np.int32)
    src = np.concatenate([shallow, shallow, deep, deep[:70], shallow[:20]])
    lv_ref, pa_ref, _ = oracle.msbfs(ro, ci, src, threads=8)
    lv, pa, st = _run_device(hg, src, ro.size - 1, True, words_per_tile=2)
    assert st.tiles == 4 and st.fallback_sources == 0 and st.max_level == 119
    assert np.array_equal(lv, lv_ref) and np.array_equal(pa, pa_ref)
