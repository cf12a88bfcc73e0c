"""Ingest formats: the text pair (parallel parse) and binary CSR files give the
same device graph, and hence bit-identical tables, as in-memory arrays."""
import numpy as np
import pytest

import oracle
import xbmsbfs as xb

pytestmark = pytest.mark.gpu


def _tables(hg, src):
    return xb.XBlossomGpuEngine10(hg).MultiSourceBFS(src, mode=xb.XB_MODE_BATCHED)


def test_text_binary_and_memory_agree(tmp_path):
    g = xb.gen_rmat(15, 16, 0.57, 0.19, 0.19, seed=4)  # column file above the parallel-parse threshold
    src = xb.sample_sources(g, 100, seed=5)
    rp, cp = tmp_path / "ro.txt", tmp_path / "ci.txt"
    np.savetxt(rp, g.row_offsets, fmt="%d")
    np.savetxt(cp, g.col_indices, fmt="%d")
    ref = _tables(xb.HGraph(g.row_offsets, g.col_indices), src)
    lr, pr, _ = oracle.msbfs(g.row_offsets, g.col_indices, src, threads=8)
    assert np.array_equal(ref[0], lr) and np.array_equal(ref[1], pr)
    from_text = xb.HGraph(str(rp), str(cp))
    assert from_text.nnodes == g.n and from_text.nedges == g.m
    t = _tables(from_text, src)
    assert np.array_equal(t[0], ref[0]) and np.array_equal(t[1], ref[1])
    bp = str(tmp_path / "g.xbcsr")
    from_text.save_binary(bp)
    from_bin = xb.HGraph.from_binary(bp)
    b = _tables(from_bin, src)
    assert np.array_equal(b[0], ref[0]) and np.array_equal(b[1], ref[1])
    ro, ci, directed = xb.load_csr_binary(bp)
    assert not directed and np.array_equal(ro, g.row_offsets) and np.array_equal(ci, g.col_indices)


def test_directed_binary_round_trip(tmp_path):
    g = xb.gen_rmat_capped(13, 6000, 60_000, 0.6, 0.15, 0.15, seed=8)
    hg = xb.HGraph(g.row_offsets, g.col_indices, directed=True)
    bp = str(tmp_path / "d.xbcsr")
    hg.save_binary(bp)
    h2 = xb.HGraph.from_binary(bp)
    assert h2.directed and h2.nnodes == g.n and h2.nedges == g.m
    src = xb.sample_sources(g, 70, seed=9)
    a, b = _tables(hg, src), _tables(h2, src)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
