"""Forest mode on the GPU (xb_match_run): bit-exact against the oracle.

The engine's matching, and its phase / level / augmentation / blossom
counts, must equal the deterministic CPU restatement's entry for entry, on the
graphs the oracle is pinned on and on larger random graphs, bipartite and
general (blossoms, nested blossoms); the size must be the maximum (Edmonds or
Hopcroft-Karp).  At full size the size-independent properties hold (a valid
matching, maximum on bipartite graphs by Hopcroft-Karp, identical on repeated
runs).  The reference's own matching test asserts validity and the maximum
size (apps/xb/test/xblsm_gpu/xblsm_gpu_10_test.cu:56-74).
"""
import ctypes

import numpy as np
import pytest

import graphs as tg
import oracle
import xbmsbfs as xb

pytestmark = pytest.mark.gpu

STATS = ("matched_vertices", "augmentations", "blossom_edges", "edges_scanned", "blossoms", "blossom_vertices",
         "phases", "levels", "aug_path_edges", "expand_edges", "blossom_scan_edges", "reused_tree_vertices",
         "reset_tree_vertices")


def _same(ro, ci, init=None, label="", regrow=False):
    """GPU vs oracle in one forest mode (tree reuse, the default, or regrow)."""
    hg = xb.HGraph(ro, ci)
    got, gs = xb.match(hg, init, regrow=regrow)
    want, ws = oracle.match_forest(ro, ci, init, reuse=not regrow)
    label = f"{label} ({'regrow' if regrow else 'reuse'})"
    assert np.array_equal(got, want), f"{label}: mate differs at {np.flatnonzero(got != want)[:8]}"
    for k in STATS:
        assert getattr(gs, k) == getattr(ws, k), f"{label}: {k} {getattr(gs, k)} != {getattr(ws, k)}"
    assert gs.reused_tree_vertices + gs.reset_tree_vertices == (ro.size - 1) * gs.phases
    tg.check_matching(ro, ci, got)
    hg.close()
    return got, gs


def _graphs():
    yield "path9", tg.path(9)
    yield "path10", tg.path(10)
    yield "cycle10", tg.cycle(10)
    yield "cycle11", tg.cycle(11)
    yield "complete8", tg.complete(8)
    yield "star9", tg.star(9)
    yield "grid37x53", tg.grid(37, 53)
    yield "K5,9", tg.complete_bipartite(5, 9)
    yield "petersen", tg.petersen()
    yield "binary_tree127", tg.binary_tree(127)
    yield "two_paths", tg.two_components(7, 8)
    for name, ro, ci, _s, _l, _p, directed in tg.golden_cases():
        if not directed:
            yield name, (ro, ci)


@pytest.mark.parametrize("regrow", [False, True], ids=["reuse", "regrow"])
@pytest.mark.parametrize("name,csr", list(_graphs()), ids=[n for n, _ in _graphs()])
def test_small_graphs_equal_oracle(name, csr, regrow):
    _same(*csr, label=name, regrow=regrow)


@pytest.mark.parametrize("regrow", [False, True], ids=["reuse", "regrow"])
@pytest.mark.parametrize("seed", range(4))
def test_random_bipartite_equal_oracle_and_maximum(seed, regrow):
    nl, nr = 20_000 + 5_000 * seed, 18_000 + 7_000 * seed
    ro, ci = tg.random_bipartite(nl, nr, 1.0 + 0.7 * seed, 40 + seed)
    got, st = _same(ro, ci, label=f"bipartite seed {seed}", regrow=regrow)
    assert st.blossom_edges == 0
    assert tg.matching_size(got) == tg.matching_size(oracle.match_hk(ro, ci))


@pytest.mark.parametrize("regrow", [False, True], ids=["reuse", "regrow"])
@pytest.mark.parametrize("seed", range(6))
def test_random_general_equal_oracle_and_maximum(seed, regrow):
    n = 400 + 700 * seed
    ro, ci = tg.gnp(n, (1.2 + 0.8 * seed) / n, 70 + seed)
    got, st = _same(ro, ci, label=f"gnp seed {seed}", regrow=regrow)
    assert tg.matching_size(got) == tg.matching_size(oracle.match_edmonds(ro, ci))


@pytest.mark.parametrize("regrow", [False, True], ids=["reuse", "regrow"])
def test_many_small_general_graphs_maximum(regrow):
    """300 small graphs (sparse to dense, and nested odd cycles): GPU equals
    the oracle and Edmonds' maximum; the blossom step runs in many of them,
    and with tree reuse later phases keep trees in many of them."""
    rng = np.random.default_rng(5)
    with_blossoms = with_reuse = 0
    for seed in range(300):
        if seed % 3 == 2:
            ro, ci = tg.nested_odd_cycles(int(rng.choice([3, 5, 7])), int(rng.integers(1, 4)), seed)
        else:
            n = int(rng.integers(3, 200))
            ro, ci = tg.gnp(n, min(1.0, float(rng.uniform(0.5, 10.0)) / n), 500 + seed)
        got, st = _same(ro, ci, label=f"small graph {seed}", regrow=regrow)
        assert tg.matching_size(got) == tg.matching_size(oracle.match_edmonds(ro, ci)), f"small graph {seed}"
        with_blossoms += st.blossoms > 0
        with_reuse += st.reused_tree_vertices > 0
    assert with_blossoms > 30
    assert (with_reuse == 0) if regrow else (with_reuse > 30)


@pytest.mark.parametrize("regrow", [False, True], ids=["reuse", "regrow"])
def test_blossom_step_is_needed(regrow):
    ro, ci = tg.csr_from_adj([[1, 2], [0, 2, 3], [0, 1], [1]])
    got, st = _same(ro, ci, np.array([-1, 2, 1, -1], dtype=np.int32), label="triangle with pendant", regrow=regrow)
    assert list(got) == [2, 3, 0, 1] and st.blossoms == 1


@pytest.mark.parametrize("regrow", [False, True], ids=["reuse", "regrow"])
@pytest.mark.parametrize("seed", range(3))
def test_large_general_equal_oracle(seed, regrow):
    """200k-vertex random graphs, average degree 3-5: many blossoms per phase."""
    n = 200_000
    ro, ci = tg.gnp(n, (3.0 + seed) / n, 900 + seed)
    got, st = _same(ro, ci, label=f"large gnp {seed}", regrow=regrow)
    assert st.blossoms > 0


@pytest.mark.parametrize("regrow", [False, True], ids=["reuse", "regrow"])
def test_rmat_equal_oracle(regrow):
    g = xb.gen_rmat(16, 8, 0.57, 0.19, 0.19, seed=21)
    _same(g.row_offsets, g.col_indices, label="rmat16", regrow=regrow)


def test_initial_matching_host_and_device():
    ro, ci = tg.random_bipartite(30_000, 25_000, 2.0, 9)
    n = ro.size - 1
    init = np.full(n, -1, dtype=np.int32)
    # Greedy start: every left vertex takes its first free neighbour.
    for v in range(30_000):
        for w in ci[ro[v]:ro[v + 1]]:
            if init[w] < 0:
                init[v], init[w] = w, v
                break
    got, st = _same(ro, ci, init, label="greedy start")
    assert st.augmentations == tg.matching_size(got) - tg.matching_size(init)
    hg = xb.HGraph(ro, ci)
    d_init = xb.DeviceBuffer((n,))
    d_init.upload(init)
    d_out = xb.DeviceBuffer((n,))
    _, ds = xb.match(hg, d_init.ptr, d_out.ptr, device=True)
    assert np.array_equal(d_out.numpy(), got) and ds.augmentations == st.augmentations
    # In place on the device (init and output the same buffer).
    xb.match(hg, d_init.ptr, d_init.ptr, device=True)
    assert np.array_equal(d_init.numpy(), got)
    d_init.free()
    d_out.free()
    hg.close()


def test_find_maximum_match_shim_sentinel():
    ro, ci = tg.grid(41, 41)  # 1681 vertices, bipartite: one exposed
    n = ro.size - 1
    eng = xb.XBlossomGpuEngine10(xb.HGraph(ro, ci))
    m = eng.FindMaximumMatch(np.full(n, n, dtype=np.int32))  # the reference runner's initial matching
    assert np.count_nonzero(m == n) == 1
    mm = m.copy()
    mm[mm == n] = -1
    tg.check_matching(ro, ci, mm)
    assert eng.last_match_stats.blossom_edges == 0
    assert np.array_equal(eng.FindMaximumMatch(), m)


def test_engine_statistics_api():
    """The counters and timing getters of xblossom_gpu_engine_10.h:39-95 on
    the Python engine: equal to the oracle's counts, adding up over calls."""
    ro, ci = tg.gnp(5_000, 3.0 / 5_000, 31)
    n = ro.size - 1
    _, ws = oracle.match_forest(ro, ci)
    eng = xb.XBlossomGpuEngine10(xb.HGraph(ro, ci), 80)
    for call in (1, 2):
        eng.FindMaximumMatch(np.full(n, n, dtype=np.int32))
        assert eng.NumBlossoms() == call * ws.blossoms
        assert eng.NumAugmentingPathEdgesProcessed() == call * ws.aug_path_edges
        assert eng.NumExpandEdgesProcessed() == call * ws.expand_edges
        assert eng.NumBlossomEdgesProcessed() == call * ws.blossom_scan_edges
        assert eng.NumEdgesProcessed() == call * ws.edges_scanned
        assert eng.NumReusedTreeNodes() == call * ws.reused_tree_vertices > 0
        assert eng.NumResetTreeNodes() == call * ws.reset_tree_vertices
        a, e, b, t = eng.GetTimingSummary()
        assert a > 0 and e > 0 and t == a + e + b
    eng.PrintTimingSummary()


def test_rejections():
    ro, ci = tg.path(6)
    hg = xb.HGraph(ro, ci)
    for bad in ([2, -1, 0, -1, -1, -1],   # (0, 2) is not an edge
                [1, 2, -1, -1, -1, -1],   # asymmetric
                [1, 0, 7, -1, -1, -1],    # out of range
                [-3, -1, -1, -1, -1, -1]):
        with pytest.raises(xb.XbError) as ei:
            xb.match(hg, np.array(bad, dtype=np.int32))
        assert ei.value.code == xb.XB_ERR_INVALID_ARG
    lib = xb.load_engine()
    out = np.zeros(6, dtype=np.int32)
    assert lib.xb_match_run(hg.handle, None, xb._ptr(out), 8, None) == xb.XB_ERR_INVALID_ARG
    hg.close()
    d = xb.HGraph(np.array([0, 1, 1], dtype=np.int32), np.array([1], dtype=np.int32), directed=True)
    with pytest.raises(xb.XbError) as ei:
        xb.match(d)
    assert ei.value.code == xb.XB_ERR_UNSUPPORTED
    d.close()


def test_edgeless_and_tiny():
    ro = np.zeros(6, dtype=np.int32)
    ci = np.zeros(0, dtype=np.int32)
    got, st = _same(ro, ci, label="edgeless")
    assert np.all(got == -1) and st.phases == 1
    _same(*tg.path(2), label="one edge")


def test_full_size_bipartite_properties():
    """1.1M vertices, 3.6M edges: valid, maximum (Hopcroft-Karp), identical on reruns."""
    ro, ci = tg.random_bipartite(600_000, 500_000, 3.0, 1)
    hg = xb.HGraph(ro, ci)
    a, st = xb.match(hg)
    b, _ = xb.match(hg)
    assert np.array_equal(a, b)
    tg.check_matching(ro, ci, a)
    assert st.blossom_edges == 0
    assert tg.matching_size(a) == tg.matching_size(oracle.match_hk(ro, ci))
    want, ws = oracle.match_forest(ro, ci)
    assert np.array_equal(a, want) and st.levels == ws.levels
    hg.close()


def test_self_loops_and_duplicate_entries():
    for seed in range(10):
        n = 500 + 300 * seed
        ro, ci = tg.gnp(n, 3.0 / n, 4000 + seed)
        ro2, ci2 = tg.with_loops_and_duplicates(ro, ci, seed)
        got, _ = _same(ro2, ci2, label=f"loops and duplicates {seed}")
        assert tg.matching_size(got) == tg.matching_size(oracle.match_edmonds(ro, ci))


def test_repeated_runs_reuse_state():
    """The graph keeps its matching state between runs: a run after a run
    from another start, and after a blossom-heavy run, is unchanged."""
    ro, ci = tg.gnp(20_000, 4.0 / 20_000, 77)
    hg = xb.HGraph(ro, ci)
    want, _ = oracle.match_forest(ro, ci)
    first, _ = xb.match(hg)
    init = np.full(ro.size - 1, -1, dtype=np.int32)
    init[0:2] = [1, 0] if 1 in ci[ro[0]:ro[1]] else [-1, -1]
    xb.match(hg, init)
    again, _ = xb.match(hg)
    assert np.array_equal(first, want) and np.array_equal(again, want)
    hg.close()


def test_reference_known_flip():
    """apps/xb/test/device/path_table_test.cu:80-124 (matching after
    AddAugmentingPath(6, 7))."""
    from test_match_oracle import KNOWN_FLIP_ADJ, KNOWN_FLIP_INIT, KNOWN_FLIP_RESULT

    ro, ci = tg.csr_from_adj(KNOWN_FLIP_ADJ)
    got, st = _same(ro, ci, np.array(KNOWN_FLIP_INIT, dtype=np.int32), label="reference known flip")
    assert list(got) == KNOWN_FLIP_RESULT


def test_general_graphs_from_greedy_starts():
    """Random greedy initial matchings on blossom-forming graphs: GPU equals
    the oracle from the same start, and reaches Edmonds' maximum."""
    rng = np.random.default_rng(12)
    for seed in range(60):
        if seed % 2:
            ro, ci = tg.nested_odd_cycles(int(rng.choice([3, 5])), int(rng.integers(1, 4)), 700 + seed)
        else:
            n = int(rng.integers(20, 400))
            ro, ci = tg.gnp(n, min(1.0, float(rng.uniform(1.0, 6.0)) / n), 800 + seed)
        n = ro.size - 1
        init = np.full(n, -1, dtype=np.int32)
        for v in rng.permutation(n):
            if init[v] >= 0:
                continue
            for w in ci[ro[v]:ro[v + 1]]:
                if init[w] < 0 and w != v:
                    init[v], init[w] = w, v
                    break
        got, st = _same(ro, ci, init, label=f"greedy start {seed}")
        best = tg.matching_size(oracle.match_edmonds(ro, ci))
        assert tg.matching_size(got) == best, f"greedy start {seed}"
        assert st.augmentations == best - tg.matching_size(init)
