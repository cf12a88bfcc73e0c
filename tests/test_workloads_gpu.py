"""Full-size named workloads on the GPU.

* R-MAT scale 16 (64 sources) and the Amazon-shaped graph (256 sources):
  every level and parent entry against the oracle;
* the LiveJournal-shaped graph (n = 4.85M, m ~ 69M, 1,024 sources, ~40 GB of
  output kept in HBM): batched tables against the plain per-source engine on
  a slice of sources, the oracle on sampled rows, the definition invariants on
  sampled rows, and a checksum of the traversed-edge counts.
"""
import numpy as np
import pytest

import graphs
import oracle
import xbmsbfs as xb

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["rmat_s16", "amazon"])
def test_full_parity(which):
    g, src = xb.workload_rmat_s16() if which == "rmat_s16" else xb.workload_amazon_like()
    lv_ref, pa_ref, e_ref = oracle.msbfs(g.row_offsets, g.col_indices, src, threads=16)
    eng = xb.XBlossomGpuEngine10(xb.HGraph(g.row_offsets, g.col_indices))
    lv, pa = eng.MultiSourceBFS(src, mode=xb.XB_MODE_BATCHED)
    assert np.array_equal(lv, lv_ref) and np.array_equal(pa, pa_ref)
    lv2, pa2 = eng.MultiSourceBFS(src[:16], mode=xb.XB_MODE_PLAIN)
    assert np.array_equal(lv2, lv_ref[:16]) and np.array_equal(pa2, pa_ref[:16])


def test_livejournal_shaped_sampled():
    g, src = xb.workload_livejournal_like()
    S, n = src.size, g.n
    eng = xb.XBlossomGpuEngine10(xb.HGraph(g.row_offsets, g.col_indices))
    lv = xb.DeviceBuffer((S, n))
    pa = xb.DeviceBuffer((S, n))
    eng.MultiSourceBFS(src, lv.ptr, pa.ptr, mode=xb.XB_MODE_BATCHED, device_out=True)
    edges = eng.CountTraversed(lv.ptr, S)
    # Plain engine on slices of sources in different words and both tiles.
    for r0, r1 in ((0, 12), (60, 70), (600, 606)):
        lvp = xb.DeviceBuffer((r1 - r0, n))
        pap = xb.DeviceBuffer((r1 - r0, n))
        eng.MultiSourceBFS(src[r0:r1], lvp.ptr, pap.ptr, mode=xb.XB_MODE_PLAIN, device_out=True)
        assert np.array_equal(lv.rows(r0, r1), lvp.numpy()) and np.array_equal(pa.rows(r0, r1), pap.numpy())
        lvp.free()
        pap.free()
    # Oracle and invariants on sampled rows.
    rows = [0, 63, 64, 511, 512, 1023]
    lv_h = np.concatenate([lv.rows(r, r + 1) for r in rows])
    pa_h = np.concatenate([pa.rows(r, r + 1) for r in rows])
    lv_ref, pa_ref, e_rows = oracle.msbfs(g.row_offsets, g.col_indices, src[rows], threads=6)
    assert np.array_equal(lv_h, lv_ref) and np.array_equal(pa_h, pa_ref)
    graphs.check_invariants(g.row_offsets, g.col_indices, src[rows], lv_h, pa_h, rows=[0, 3])
    assert 0 < edges <= S * g.m
    deg = g.degrees().astype(np.int64)
    assert e_rows == int(sum(deg[lv_h[i] >= 0].sum() for i in range(len(rows))))
