"""CPU tests that pin the oracle (no GPU needed).

* networkx golden fixtures (third-party shortest-path lengths and predecessor
  sets, committed under tests/golden/),
* closed forms on paths, cycles, stars, complete graphs, grids and binary
  trees,
* the reference's own known-answer tests for this path:
  - edge-balanced mapping, apps/xb/test/device/scan_bs_test.cu:58-137
    (degrees {5,4,3,2,1}: 15 tasklets, tasklet -> (node, edge offset));
  - path-table root distances, apps/xb/test/device/path_table_test.cu:80-124
    and :130-156 (ExpandTwoNodes trees: root_distance = hop distance),
* the bit-parallel restatement against the per-source one,
* planted mistakes in oracle outputs, which the invariant checker must catch.
"""
import numpy as np
import pytest

import graphs
import oracle


@pytest.mark.parametrize("case", list(graphs.golden_cases()), ids=lambda c: c[0])
def test_oracle_matches_networkx(case):
    name, ro, ci, src, lv_ref, pa_ref, directed = case
    lv, pa, _ = oracle.msbfs(ro, ci, src, threads=2)
    assert np.array_equal(lv, lv_ref), name
    assert np.array_equal(pa, pa_ref), name


@pytest.mark.parametrize("case", list(graphs.closed_form_cases()), ids=lambda c: c[0])
def test_oracle_closed_forms(case):
    name, ro, ci, src, lv_ref, pa_ref = case
    lv, pa, _ = oracle.msbfs(ro, ci, src)
    assert np.array_equal(lv, lv_ref), name
    assert np.array_equal(pa, pa_ref), name
    graphs.check_invariants(ro, ci, src, lv, pa)


def test_disconnected_components():
    ro, ci = graphs.two_components(10, 7)
    lv, pa, e = oracle.msbfs(ro, ci, np.array([3, 12], dtype=np.int32))
    assert np.all(lv[0, 10:] == -1) and np.all(pa[0, 10:] == -1)
    assert np.all(lv[1, :10] == -1) and np.all(pa[1, :10] == -1)
    assert lv[0, 9] == 6 and lv[1, 16] == 4 and lv[1, 10] == 2
    assert e == int(np.diff(ro)[:10].sum() + np.diff(ro)[10:].sum())


def test_edge_map_known_answer_scan_bs():
    # apps/xb/test/device/scan_bs_test.cu:58-137: node ids 1..5 with degrees 5,4,3,2,1.
    deg = np.array([5, 4, 3, 2, 1], dtype=np.int32)
    presum = oracle.exclusive_scan(deg)
    assert presum.tolist() == [0, 5, 9, 12, 14, 15]
    owner, offset = oracle.edge_map(deg)
    assert owner.size == 15
    node_ids = np.array([1, 2, 3, 4, 5])
    exp_nid, exp_eid = [], []
    for i, d in enumerate(deg):
        exp_nid += [node_ids[i]] * int(d)
        exp_eid += list(range(int(d)))
    assert node_ids[owner].tolist() == exp_nid
    assert offset.tolist() == exp_eid


def test_edge_map_skips_empty_segments():
    deg = np.array([0, 3, 0, 0, 2, 0], dtype=np.int32)
    owner, offset = oracle.edge_map(deg)
    assert owner.tolist() == [1, 1, 1, 4, 4]
    assert offset.tolist() == [0, 1, 2, 0, 1]


def _expand_two_nodes_graph(n, triples):
    """Edges of an ExpandTwoNodes(v, w, x) tree: v-w and w-x."""
    adj = [[] for _ in range(n)]
    for v, w, x in triples:
        adj[v].append(w)
        adj[w].append(v)
        adj[w].append(x)
        adj[x].append(w)
    return graphs.csr_from_adj(adj)


def test_path_table_root_distance_graph1():
    # apps/xb/test/device/path_table_test.cu:80-100: ExpandTwoNodes(0,1,2),(0,3,4),(4,5,6)
    # gives root_distance {0:0, 2:2, 4:2, 6:4}; the recorded pair for x=6 is [w=5, v=4].
    ro, ci = _expand_two_nodes_graph(8, [(0, 1, 2), (0, 3, 4), (4, 5, 6)])
    lv, pa, _ = oracle.msbfs(ro, ci, np.array([0], dtype=np.int32))
    assert {v: int(lv[0, v]) for v in (0, 2, 4, 6)} == {0: 0, 2: 2, 4: 2, 6: 4}
    assert pa[0, 6] == 5 and pa[0, 5] == 4 and pa[0, 2] == 1 and pa[0, 1] == 0
    assert lv[0, 7] == -1


def test_path_table_root_distance_graph2():
    # apps/xb/test/device/path_table_test.cu:130-156 (19 vertices).
    triples = [(0, 1, 2), (2, 3, 4), (2, 5, 6), (4, 7, 8), (8, 9, 10), (8, 11, 12), (6, 13, 14), (6, 15, 16),
               (16, 17, 18)]
    ro, ci = _expand_two_nodes_graph(19, triples)
    lv, _, _ = oracle.msbfs(ro, ci, np.array([0], dtype=np.int32))
    exp = {0: 0, 2: 2, 4: 4, 6: 4, 8: 6, 10: 8, 12: 8, 18: 8}
    assert {v: int(lv[0, v]) for v in exp} == exp


@pytest.mark.parametrize("case", list(graphs.golden_cases()), ids=lambda c: c[0])
def test_bitparallel_oracle_agrees(case):
    name, ro, ci, src, lv_ref, pa_ref, directed = case
    lv, pa, e = oracle.msbfs(ro, ci, src, bitparallel=True)
    _, _, e1 = oracle.msbfs(ro, ci, src)
    assert np.array_equal(lv, lv_ref) and np.array_equal(pa, pa_ref) and e == e1


def test_invariant_checker_catches_planted_mistakes():
    """Plausible oracle bugs, planted in oracle output, must fail the checker."""
    name, ro, ci, src, lv, pa, directed = next(c for c in graphs.golden_cases() if c[0] == "gnp200_dense")
    graphs.check_invariants(ro, ci, src, lv, pa)
    n = ro.size - 1
    deg = np.diff(ro)
    # 1) highest-id instead of lowest-id predecessor
    pa_hi = pa.copy()
    r = 0
    L = lv[r]
    changed = False
    for v in range(n):
        if L[v] >= 1:
            preds = [u for u in range(n) if v in ci[ro[u]:ro[u + 1]] and L[u] == L[v] - 1]
            if len(preds) > 1:
                pa_hi[r, v] = max(preds)
                changed = True
                break
    assert changed
    with pytest.raises(AssertionError):
        graphs.check_invariants(ro, ci, src, lv, pa_hi, rows=[r])
    # 2) off-by-one level
    lv_bad = lv.copy()
    v = int(np.flatnonzero(lv[1] == 2)[0])
    lv_bad[1, v] = 3
    with pytest.raises(AssertionError):
        graphs.check_invariants(ro, ci, src, lv_bad, pa, rows=[1])
    # 3) parent taken one level too far (a grand-parent)
    pa_gp = pa.copy()
    v = int(np.flatnonzero(lv[2] == 2)[0])
    pa_gp[2, v] = pa[2, pa[2, v]]
    with pytest.raises(AssertionError):
        graphs.check_invariants(ro, ci, src, lv, pa_gp, rows=[2])
    # 4) transposed edges on a directed fixture (out- instead of in-neighbours)
    name, dro, dci, dsrc, dlv, dpa, _ = next(c for c in graphs.golden_cases() if c[0] == "digraph250")
    m = dci.size
    srcs = np.repeat(np.arange(dro.size - 1), np.diff(dro))
    order = np.lexsort((srcs, dci))
    tro = np.zeros(dro.size, dtype=np.int32)
    np.add.at(tro, dci + 1, 1)
    tro = np.cumsum(tro).astype(np.int32)
    tci = srcs[order].astype(np.int32)
    tlv, tpa, _ = oracle.msbfs(tro, tci, dsrc)
    assert not (np.array_equal(tlv, dlv) and np.array_equal(tpa, dpa))
    assert deg.sum() > 0 and m > 0
