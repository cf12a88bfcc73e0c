"""GPU parity: the HIP engine through the C ABI against the CPU oracle.

Bar: bit-exact levels and parents (integer work).  Cases cover both modes,
every words-per-tile value, forced push-only / pull-only direction choices,
heavy-vertex kernels (tiny cuts), directed graphs, partial tiles, duplicate
and isolated sources, raw input with self-loops and repeated entries, zero
sources, and depth beyond the 8-bit level code (plain fallback).
"""
import numpy as np
import pytest

import graphs
import oracle
import xbmsbfs as xb

pytestmark = pytest.mark.gpu


def _path_graph(n):
    ro = np.zeros(n + 1, dtype=np.int32)
    cols = []
    for v in range(n):
        nb = [u for u in (v - 1, v + 1) if 0 <= u < n]
        cols.extend(nb)
        ro[v + 1] = ro[v] + len(nb)
    return xb.Csr(ro, np.array(cols, dtype=np.int32))


@pytest.fixture(scope="module")
def rmat12():
    g = xb.gen_rmat(12, 8, 0.57, 0.19, 0.19, seed=11)
    return g, xb.HGraph(g.row_offsets, g.col_indices)


def _check(g, hg, src, **kw):
    lv_ref, pa_ref, _ = oracle.msbfs(g.row_offsets, g.col_indices, src, threads=8)
    eng = xb.XBlossomGpuEngine10(hg)
    lv, pa = eng.MultiSourceBFS(src, **kw)
    assert np.array_equal(lv, lv_ref), f"levels differ in {(lv != lv_ref).sum()} entries ({kw})"
    assert np.array_equal(pa, pa_ref), f"parents differ in {(pa != pa_ref).sum()} entries ({kw})"
    return eng.last_stats


@pytest.mark.parametrize("T", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("alpha", [1, 20, 1_000_000_000])
def test_batched_rmat(rmat12, T, alpha):
    g, hg = rmat12
    src = xb.sample_sources(g, 150, seed=3)
    st = _check(g, hg, src, mode=xb.XB_MODE_BATCHED, words_per_tile=T, alpha=alpha)
    assert st.tiles == -(-150 // (64 * T))


@pytest.mark.parametrize("T", [1, 8, 16])
def test_batched_heavy_kernels(rmat12, T):
    g, hg = rmat12
    src = xb.sample_sources(g, 200, seed=4)
    _check(g, hg, src, mode=xb.XB_MODE_BATCHED, words_per_tile=T, pull_cut_small=1, pull_cut_medium=3, pull_cut_hub=6, hub_chunk=5)


def test_plain_rmat(rmat12):
    g, hg = rmat12
    src = xb.sample_sources(g, 40, seed=5)
    _check(g, hg, src, mode=xb.XB_MODE_PLAIN)


def test_duplicate_and_isolated_sources():
    g = xb.gen_rmat(10, 4, 0.57, 0.19, 0.19, seed=7)
    deg = g.degrees()
    iso = np.flatnonzero(deg == 0)
    assert iso.size > 0
    src = np.array([iso[0], 5, 5, 17, iso[-1], 5] * 13, dtype=np.int32)
    hg = xb.HGraph(g.row_offsets, g.col_indices)
    for mode in (xb.XB_MODE_BATCHED, xb.XB_MODE_PLAIN):
        _check(g, hg, src, mode=mode)


def test_directed_rmat():
    g = xb.gen_rmat(11, 8, 0.6, 0.15, 0.15, seed=9, symmetric=False)
    hg = xb.HGraph(g.row_offsets, g.col_indices, directed=True)
    src = xb.sample_sources(g, 130, seed=1)
    for mode in (xb.XB_MODE_BATCHED, xb.XB_MODE_PLAIN):
        _check(g, hg, src, mode=mode)
    _check(g, hg, src, mode=xb.XB_MODE_BATCHED, words_per_tile=2, pull_cut_small=2, pull_cut_medium=2, pull_cut_hub=4, hub_chunk=3,
           alpha=1_000_000)


def test_chung_lu_small():
    g = xb.gen_chung_lu(20_000, 2.1, 10.0, 800, seed=3)
    hg = xb.HGraph(g.row_offsets, g.col_indices)
    src = xb.sample_sources(g, 300, seed=3, degree_weighted=True)
    _check(g, hg, src, mode=xb.XB_MODE_BATCHED, words_per_tile=4)


def test_deep_path_falls_back():
    g = _path_graph(600)
    hg = xb.HGraph(g.row_offsets, g.col_indices)
    src = np.array([0, 599, 300, 10], dtype=np.int32)
    st = _check(g, hg, src, mode=xb.XB_MODE_BATCHED)
    assert st.fallback_sources == 4
    src2 = np.array([250, 300, 350], dtype=np.int32)  # depth <= 350 -> still deep
    _check(g, hg, src2, mode=xb.XB_MODE_PLAIN)


def test_medium_path_no_fallback():
    g = _path_graph(200)
    hg = xb.HGraph(g.row_offsets, g.col_indices)
    src = np.arange(0, 200, 3, dtype=np.int32)
    st = _check(g, hg, src, mode=xb.XB_MODE_BATCHED)
    assert st.fallback_sources == 0 and st.max_level == 199


def test_deep_and_shallow_tiles_share_planes():
    """Level codes 16 and above use planes 4-7, which a tile clears only when
    it first reaches level 15.  A 120-vertex path (levels to 119) beside a
    40-leaf star (levels <= 2): 16-word tiles alternate shallow / deep /
    shallow / deep in one run and across runs on the same workspace."""
    import graphs as tg

    ro, ci = tg.two_components(120, 1)
    star_ro, star_ci = tg.star(41)
    n1 = ro.size - 1
    ro = np.concatenate([ro, star_ro[1:] + ro[-1]]).astype(np.int32)
    ci = np.concatenate([ci, star_ci + n1]).astype(np.int32)
    g = xb.Csr(ro, ci)
    hg = xb.HGraph(ro, ci)
    rng = np.random.default_rng(31)
    shallow = rng.integers(n1, n1 + 41, size=1024).astype(np.int32)
    deep = rng.integers(0, 120, size=1024).astype(np.int32)
    for order in ((shallow, deep, shallow, deep[:300]), (deep, shallow[:700], deep[:100])):
        src = np.concatenate(order)
        st = _check(g, hg, src, mode=xb.XB_MODE_BATCHED, words_per_tile=16)
        assert st.fallback_sources == 0 and st.max_level == 119


@pytest.mark.parametrize("T", [1, 2, 4, 8, 16])
def test_bucket_mix(T):
    """Every pull bucket class plus hub pieces on one graph (cut-offs chosen so all are populated)."""
    g = xb.gen_chung_lu(30_000, 2.1, 12.0, 3_000, seed=21)
    hg = xb.HGraph(g.row_offsets, g.col_indices)
    src = xb.sample_sources(g, 64 * T + 5, seed=22)
    deg = g.degrees()
    assert (deg > 200).sum() > 0 and ((deg > 4) & (deg <= 20)).sum() > 0
    _check(g, hg, src, mode=xb.XB_MODE_BATCHED, words_per_tile=T, pull_cut_small=4, pull_cut_medium=20,
           pull_cut_hub=200, hub_chunk=64, alpha=100)


# ---- pinned cases: networkx fixtures, closed forms, reference known answers ----
import graphs  # noqa: E402


@pytest.mark.parametrize("case", list(graphs.golden_cases()), ids=lambda c: c[0])
def test_golden_fixtures(case):
    name, ro, ci, src, lv_ref, pa_ref, directed = case
    hg = xb.HGraph(ro, ci, directed=directed)
    eng = xb.XBlossomGpuEngine10(hg)
    for kw in ({"mode": xb.XB_MODE_PLAIN}, {"mode": xb.XB_MODE_BATCHED, "words_per_tile": 1},
               {"mode": xb.XB_MODE_BATCHED, "words_per_tile": 8, "alpha": 1_000_000}):
        lv, pa = eng.MultiSourceBFS(src, **kw)
        assert np.array_equal(lv, lv_ref), (name, kw)
        assert np.array_equal(pa, pa_ref), (name, kw)


@pytest.mark.parametrize("case", list(graphs.closed_form_cases()), ids=lambda c: c[0])
def test_closed_forms(case):
    name, ro, ci, src, lv_ref, pa_ref = case
    hg = xb.HGraph(ro, ci)
    eng = xb.XBlossomGpuEngine10(hg)
    for mode in (xb.XB_MODE_PLAIN, xb.XB_MODE_BATCHED):
        lv, pa = eng.MultiSourceBFS(src, mode=mode)
        assert np.array_equal(lv, lv_ref) and np.array_equal(pa, pa_ref), (name, mode)


def test_path_table_known_answers():
    # apps/xb/test/device/path_table_test.cu:80-124 and :130-156: the root
    # distances of ExpandTwoNodes trees are BFS hop distances from the root.
    def tree(n, triples):
        adj = [[] for _ in range(n)]
        for v, w, x in triples:
            adj[v] += [w]
            adj[w] += [v, x]
            adj[x] += [w]
        return graphs.csr_from_adj(adj)

    ro, ci = tree(8, [(0, 1, 2), (0, 3, 4), (4, 5, 6)])
    lv, pa = xb.XBlossomGpuEngine10(xb.HGraph(ro, ci)).MultiSourceBFS([0], mode=xb.XB_MODE_BATCHED)
    assert [int(lv[0, v]) for v in (0, 2, 4, 6)] == [0, 2, 2, 4] and pa[0, 6] == 5 and lv[0, 7] == -1
    ro, ci = tree(19, [(0, 1, 2), (2, 3, 4), (2, 5, 6), (4, 7, 8), (8, 9, 10), (8, 11, 12), (6, 13, 14),
                       (6, 15, 16), (16, 17, 18)])
    lv, _ = xb.XBlossomGpuEngine10(xb.HGraph(ro, ci)).MultiSourceBFS([0], mode=xb.XB_MODE_BATCHED)
    exp = {0: 0, 2: 2, 4: 4, 6: 4, 8: 6, 10: 8, 12: 8, 18: 8}
    assert {v: int(lv[0, v]) for v in exp} == exp


def _pairs(owner, offset):
    return sorted(zip(owner.tolist(), offset.tolist()))


def test_device_edge_map_known_answer():
    # apps/xb/test/device/scan_bs_test.cu:58-137: degrees {5,4,3,2,1} -> 15 tasklets.
    deg = np.array([5, 4, 3, 2, 1], dtype=np.int32)
    owner, offset = xb.test_edge_map(deg)
    ro, oo = oracle.edge_map(deg)
    assert owner.size == 15 and _pairs(owner, offset) == _pairs(ro, oo)


@pytest.mark.parametrize("seed", [0, 1])
def test_device_edge_map_large_and_zero_runs(seed):
    rng = np.random.default_rng(seed)
    deg = rng.integers(0, 40, size=50_000).astype(np.int32)
    deg[rng.random(deg.size) < 0.7] = 0  # long runs of empty segments force the global-search path
    deg[123] = 5000
    owner, offset = xb.test_edge_map(deg)
    ro, oo = oracle.edge_map(deg)
    assert _pairs(owner, offset) == _pairs(ro, oo)


def test_runs_are_bit_identical(rmat12):
    """Deterministic lowest-id rule: repeated runs give identical tables (race check)."""
    g, hg = rmat12
    src = xb.sample_sources(g, 300, seed=9)
    eng = xb.XBlossomGpuEngine10(hg)
    a = eng.MultiSourceBFS(src, mode=xb.XB_MODE_BATCHED, words_per_tile=4)
    b = eng.MultiSourceBFS(src, mode=xb.XB_MODE_BATCHED, words_per_tile=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_frontier_words_left_between_tiles_and_runs(rmat12):
    """Levels leave their frontier words set; the next tile clears both
    frontier arrays over the words the last tile used.  One engine (one
    workspace) runs tiles of 16, 2 and 1 words and back again, with partial
    last tiles, pull-only and push-only direction choices, and each run's
    tables equal the oracle's: a left word may not leak into the next tile or
    run, whatever its width."""
    g, hg = rmat12
    eng = xb.XBlossomGpuEngine10(hg)
    for k, (S, T, alpha) in enumerate([(1100, 16, 20), (130, 2, 1), (40, 1, 1_000_000_000),
                                       (1024, 16, 1), (70, 16, 20), (200, 4, 20)]):
        src = xb.sample_sources(g, S, seed=40 + k)
        lv_ref, pa_ref, _ = oracle.msbfs(g.row_offsets, g.col_indices, src, threads=8)
        lv, pa = eng.MultiSourceBFS(src, mode=xb.XB_MODE_BATCHED, words_per_tile=T, alpha=alpha)
        assert np.array_equal(lv, lv_ref), f"run {k}: levels differ in {(lv != lv_ref).sum()} entries"
        assert np.array_equal(pa, pa_ref), f"run {k}: parents differ in {(pa != pa_ref).sum()} entries"


def test_source_out_of_range_rejected(rmat12):
    g, hg = rmat12
    eng = xb.XBlossomGpuEngine10(hg)
    with pytest.raises(xb.XbError) as ei:
        eng.MultiSourceBFS([0, g.n], mode=xb.XB_MODE_BATCHED)
    assert ei.value.code == xb.XB_ERR_SOURCE_RANGE
    with pytest.raises(xb.XbError) as ei:
        eng.MultiSourceBFS([-1], mode=xb.XB_MODE_PLAIN)
    assert ei.value.code == xb.XB_ERR_SOURCE_RANGE


def test_text_loader_roundtrip(tmp_path, rmat12):
    g, _ = rmat12
    (tmp_path / "ro.txt").write_text(" ".join(map(str, g.row_offsets.tolist())) + "\n")
    (tmp_path / "ci.txt").write_text("\n".join(map(str, g.col_indices.tolist())) + "\n")
    hg = xb.HGraph(str(tmp_path / "ro.txt"), str(tmp_path / "ci.txt"))
    assert (hg.nnodes, hg.nedges) == (g.n, g.m)
    _check(g, hg, xb.sample_sources(g, 70, seed=1), mode=xb.XB_MODE_BATCHED)


def test_unsorted_adjacency_is_sorted_on_ingest(rmat12):
    g, _ = rmat12
    ci = g.col_indices.copy()
    for v in range(0, g.n, 7):
        a, b = g.row_offsets[v], g.row_offsets[v + 1]
        ci[a:b] = ci[a:b][::-1]
    hg = xb.HGraph(g.row_offsets, ci)
    _check(g, hg, xb.sample_sources(g, 100, seed=2), mode=xb.XB_MODE_BATCHED)


@pytest.mark.parametrize("T", [1, 4, 16])
def test_self_loops_and_repeated_entries(rmat12, T):
    """Raw CSR input may hold self-loops and repeated entries: HGraph takes
    the files as they are (apps/xb/src/include/graph/hgraph_gpu.h:37-48).
    Neither changes a level or a lowest-id parent: the tables equal the
    oracle's on the raw input and the oracle's on the clean graph.  Tiny
    bucket cuts send the repeated entries through the hub-piece pull too."""
    g, _ = rmat12
    ro, ci = graphs.with_loops_and_duplicates(g.row_offsets, g.col_indices, seed=9)
    assert ci.size > g.m
    raw = xb.Csr(ro, ci)
    hg = xb.HGraph(ro, ci)
    src = xb.sample_sources(g, 130, seed=6)
    lv_clean, pa_clean, _ = oracle.msbfs(g.row_offsets, g.col_indices, src, threads=8)
    for kw in ({}, dict(pull_cut_small=1, pull_cut_medium=3, pull_cut_hub=6, hub_chunk=5)):
        _check(raw, hg, src, mode=xb.XB_MODE_BATCHED, words_per_tile=T, **kw)
        lv, pa = xb.XBlossomGpuEngine10(hg).MultiSourceBFS(src, mode=xb.XB_MODE_BATCHED, words_per_tile=T, **kw)
        assert np.array_equal(lv, lv_clean) and np.array_equal(pa, pa_clean)
    _check(raw, hg, src[:20], mode=xb.XB_MODE_PLAIN)


def test_zero_sources(rmat12):
    """S = 0 is a valid call: nothing is written and the stats read zero."""
    _, hg = rmat12
    eng = xb.XBlossomGpuEngine10(hg)
    eng.MultiSourceBFS(xb.sample_sources(rmat12[0], 10, seed=1), mode=xb.XB_MODE_BATCHED)
    assert eng.last_stats.ms_total > 0
    lv, pa = eng.MultiSourceBFS(np.zeros(0, dtype=np.int32), mode=xb.XB_MODE_BATCHED)
    assert lv.shape == (0, hg.nnodes) and pa.shape == (0, hg.nnodes)
    assert eng.last_stats.ms_total == 0 and eng.last_stats.tiles == 0


def test_cpp_shim_call_site(tmp_path):
    """The C++ shim (zblossom::HGraph / XBlossomGpuEngine10) from a C++ call site."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(xb.ENGINE_LIB), "shim_test")
    assert os.path.exists(exe), "shim_test not built"
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "shim_test ok" in r.stdout
