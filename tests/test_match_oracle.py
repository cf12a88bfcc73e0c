"""Pins the matching oracle (oracle/match_oracle.c) before it checks the engine.

* Edmonds' algorithm: closed forms (paths, cycles, complete graphs, stars,
  grids, complete bipartite graphs, the Petersen graph) and networkx's maximum
  cardinality matching on random graphs.
* Hopcroft-Karp: closed forms, scipy's maximum_bipartite_matching, and the
  refusal of non-bipartite input.
* The deterministic forest restatement with the blossom step (the engine's
  checker): valid matchings everywhere, and maximum everywhere -- against
  closed forms, Hopcroft-Karp on bipartite graphs, and Edmonds on random,
  dense and nested-blossom general graphs, from empty and greedy starts.
The reference's own matching tests assert validity and the maximum size on
datasets that are not shipped (apps/xb/test/xblsm_gpu/xblsm_gpu_10_test.cu:
56-74); the closed forms and library results stand in for those sizes.
"""
import numpy as np
import pytest

import graphs as tg
import oracle

CLOSED = {
    "path9": (lambda: tg.path(9), 4, True),
    "path10": (lambda: tg.path(10), 5, True),
    "cycle10": (lambda: tg.cycle(10), 5, True),
    "cycle11": (lambda: tg.cycle(11), 5, False),
    "complete7": (lambda: tg.complete(7), 3, False),
    "complete8": (lambda: tg.complete(8), 4, False),
    "star9": (lambda: tg.star(9), 1, True),
    "grid5x7": (lambda: tg.grid(5, 7), 17, True),
    "grid6x6": (lambda: tg.grid(6, 6), 18, True),
    "K3,8": (lambda: tg.complete_bipartite(3, 8), 3, True),
    "K6,6": (lambda: tg.complete_bipartite(6, 6), 6, True),
    "petersen": (lambda: tg.petersen(), 5, False),
    "binary_tree15": (lambda: tg.binary_tree(15), 5, True),  # the 5 even-depth vertices cover every edge
    "two_paths": (lambda: tg.two_components(5, 6), 5, True),
}


@pytest.mark.parametrize("name", sorted(CLOSED))
def test_edmonds_closed_forms(name):
    mk, size, _ = CLOSED[name]
    ro, ci = mk()
    mate = oracle.match_edmonds(ro, ci)
    tg.check_matching(ro, ci, mate)
    assert tg.matching_size(mate) == size


@pytest.mark.parametrize("name", sorted(CLOSED))
def test_hk_closed_forms(name):
    mk, size, bip = CLOSED[name]
    ro, ci = mk()
    mate = oracle.match_hk(ro, ci)
    if not bip:
        assert mate is None
        return
    tg.check_matching(ro, ci, mate)
    assert tg.matching_size(mate) == size


@pytest.mark.parametrize("name", sorted(CLOSED))
def test_forest_closed_forms(name):
    mk, size, bip = CLOSED[name]
    ro, ci = mk()
    mate, st = oracle.match_forest(ro, ci)
    tg.check_matching(ro, ci, mate)
    got = tg.matching_size(mate)
    assert st.matched_vertices == 2 * got
    assert got == size
    if bip:
        assert st.blossoms == 0 and st.blossom_edges == 0


@pytest.mark.parametrize("seed", range(6))
def test_edmonds_against_networkx(seed):
    nx = pytest.importorskip("networkx")
    n = 40 + 30 * seed
    ro, ci = tg.gnp(n, 2.5 / n, seed)
    mate = oracle.match_edmonds(ro, ci)
    tg.check_matching(ro, ci, mate)
    G = nx.Graph()
    G.add_nodes_from(range(n))
    for v in range(n):
        G.add_edges_from((v, int(w)) for w in ci[ro[v]:ro[v + 1]])
    assert tg.matching_size(mate) == len(nx.max_weight_matching(G, maxcardinality=True))


@pytest.mark.parametrize("seed", range(6))
def test_hk_against_scipy(seed):
    sp = pytest.importorskip("scipy.sparse")
    csgraph = pytest.importorskip("scipy.sparse.csgraph")
    nl, nr = 300 + 50 * seed, 250 + 70 * seed
    ro, ci = tg.random_bipartite(nl, nr, 1.0 + 0.5 * seed, seed)
    mate = oracle.match_hk(ro, ci)
    tg.check_matching(ro, ci, mate)
    rows = np.repeat(np.arange(nl), np.diff(ro)[:nl])
    cols = ci[:ro[nl]] - nl
    bi = sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(nl, nr))
    ref = csgraph.maximum_bipartite_matching(bi, perm_type="column")
    assert tg.matching_size(mate) == int(np.count_nonzero(ref >= 0))


@pytest.mark.parametrize("reuse", [True, False], ids=["reuse", "regrow"])
@pytest.mark.parametrize("seed", range(8))
def test_forest_maximum_on_bipartite(seed, reuse):
    nl, nr = 200 + 40 * seed, 220 + 30 * seed
    ro, ci = tg.random_bipartite(nl, nr, 0.8 + 0.4 * seed, 100 + seed)
    mate, st = oracle.match_forest(ro, ci, reuse=reuse)
    tg.check_matching(ro, ci, mate)
    assert st.blossom_edges == 0
    assert tg.matching_size(mate) == tg.matching_size(oracle.match_hk(ro, ci))
    assert st.augmentations == tg.matching_size(mate)


def _greedy(ro, ci, order):
    n = ro.size - 1
    init = np.full(n, -1, dtype=np.int32)
    for v in order:
        if init[v] >= 0:
            continue
        for w in ci[ro[v]:ro[v + 1]]:
            if init[w] < 0 and w != v:
                init[v], init[w] = w, v
                break
    return init


@pytest.mark.parametrize("reuse", [True, False], ids=["reuse", "regrow"])
def test_forest_maximum_on_general_graphs(reuse):
    """Equal to Edmonds' maximum on sparse, dense and nested-blossom graphs,
    from the empty matching and from a greedy start in random order, whether
    the trees of still-exposed roots are kept between phases or not."""
    rng = np.random.default_rng(0)
    with_blossoms = with_reuse = 0
    for seed in range(400):
        if seed % 4 == 3:
            ro, ci = tg.nested_odd_cycles(int(rng.choice([3, 5, 7])), int(rng.integers(1, 4)), seed)
        else:
            n = int(rng.integers(3, 160))
            ro, ci = tg.gnp(n, min(1.0, float(rng.uniform(0.5, 8.0)) / n), 1000 + seed)
        best = tg.matching_size(oracle.match_edmonds(ro, ci))
        mate, st = oracle.match_forest(ro, ci, reuse=reuse)
        tg.check_matching(ro, ci, mate)
        assert tg.matching_size(mate) == best, f"graph {seed}"
        with_blossoms += st.blossoms > 0
        with_reuse += st.reused_tree_vertices > 0
        if seed % 2 == 0:
            init = _greedy(ro, ci, rng.permutation(ro.size - 1))
            mate2, st2 = oracle.match_forest(ro, ci, init, reuse=reuse)
            tg.check_matching(ro, ci, mate2)
            assert tg.matching_size(mate2) == best, f"graph {seed}, greedy start"
            assert st2.augmentations == best - tg.matching_size(init)
    assert with_blossoms > 50  # the blossom step is exercised, not just bipartite-like luck
    assert (with_reuse > 50) if reuse else (with_reuse == 0)


def test_tree_reuse_counters_and_levels():
    """Every phase set-up counts each vertex once (reused or reset), the first
    phase resets all of them, regrowing never reuses, and keeping trees grows
    fewer levels in total (the point of the reference's lazy reset)."""
    rng = np.random.default_rng(3)
    levels = {True: 0, False: 0}
    for seed in range(120):
        n = int(rng.integers(50, 400))
        ro, ci = tg.gnp(n, float(rng.uniform(1.0, 6.0)) / n, 2000 + seed)
        for reuse in (True, False):
            _, st = oracle.match_forest(ro, ci, reuse=reuse)
            assert st.reused_tree_vertices + st.reset_tree_vertices == n * st.phases
            assert st.edges_scanned == st.aug_path_edges + st.expand_edges + st.blossom_scan_edges
            if not reuse:
                assert st.reused_tree_vertices == 0
            _, st1 = oracle.match_forest(ro, ci, reuse=reuse, max_phases=1)
            assert st1.reset_tree_vertices == n and st1.reused_tree_vertices == 0
            levels[reuse] += st.levels
    assert levels[True] < levels[False]


def test_blossom_step_is_needed():
    """Triangle {0, 1, 2} with 1-2 matched, pendant 3 on vertex 1.  The forest
    from roots 0 and 3 makes 1 odd and 2 even (tree 0); 2's only even
    neighbour is 0 in its own tree, so without the blossom step the phase
    ends with one pair.  The blossom {0, 1, 2} makes 1 outer, and its edge to
    root 3 closes the augmenting path 3-1=2-0: two pairs."""
    ro, ci = tg.csr_from_adj([[1, 2], [0, 2, 3], [0, 1], [1]])
    init = np.array([-1, 2, 1, -1], dtype=np.int32)
    mate, st = oracle.match_forest(ro, ci, init)
    tg.check_matching(ro, ci, mate)
    assert tg.matching_size(mate) == 2
    assert st.blossoms == 1 and st.blossom_vertices == 1 and st.augmentations == 1
    assert list(mate) == [2, 3, 0, 1]


def test_forest_from_an_initial_matching():
    ro, ci = tg.random_bipartite(500, 450, 2.0, 7)
    n = ro.size - 1
    # Greedy initial matching in vertex order.
    init = _greedy(ro, ci, range(n))
    tg.check_matching(ro, ci, init)
    mate, st = oracle.match_forest(ro, ci, init)
    tg.check_matching(ro, ci, mate)
    assert tg.matching_size(mate) == tg.matching_size(oracle.match_hk(ro, ci))
    assert st.augmentations == tg.matching_size(mate) - tg.matching_size(init)
    # A maximum matching as the start: one phase, nothing to flip.
    again, st2 = oracle.match_forest(ro, ci, mate)
    assert np.array_equal(again, mate) and st2.phases == 1 and st2.augmentations == 0


def test_violation_counter():
    ro, ci = tg.path(6)
    good = np.array([1, 0, 3, 2, 5, 4], dtype=np.int32)
    assert oracle.match_violations(ro, ci, good) == 0
    assert oracle.match_violations(ro, ci, np.array([2, -1, 0, -1, -1, -1], dtype=np.int32)) == 2  # not an edge
    assert oracle.match_violations(ro, ci, np.array([1, 2, -1, -1, -1, -1], dtype=np.int32)) == 2  # asymmetric
    assert oracle.match_violations(ro, ci, np.array([-2, -1, -1, -1, -1, -1], dtype=np.int32)) == 1


def test_edge_cases():
    for ro, ci in [(np.zeros(1, np.int32), np.zeros(0, np.int32)),
                   (np.zeros(5, np.int32), np.zeros(0, np.int32)),
                   tg.path(2), tg.path(3)]:
        mate, st = oracle.match_forest(ro, ci)
        tg.check_matching(ro, ci, mate)
        assert tg.matching_size(mate) == tg.matching_size(oracle.match_edmonds(ro, ci))


def test_self_loops_and_duplicate_entries():
    """Loops never match and repeated entries change nothing: the size equals
    Edmonds' maximum on the clean graph."""
    for seed in range(40):
        n = 20 + 5 * seed
        ro, ci = tg.gnp(n, 3.0 / n, 3000 + seed)
        ro2, ci2 = tg.with_loops_and_duplicates(ro, ci, seed)
        mate, _ = oracle.match_forest(ro2, ci2)
        assert oracle.match_violations(ro, ci, mate) == 0
        assert tg.matching_size(mate) == tg.matching_size(oracle.match_edmonds(ro, ci))


# The reference's known answer for a path flip
# (apps/xb/test/device/path_table_test.cu:80-124): tree 0 -(1=2), 0 -(3=4),
# 4 -(5=6) with 7 exposed; the augmenting path 0-3=4-5=6-7 flips to the
# matching below (1=2 untouched).  As a graph with that initial matching, the
# forest finds the same (only) augmenting path.
KNOWN_FLIP_ADJ = [[1, 3], [0, 2], [1], [0, 4], [3, 5], [4, 6], [5, 7], [6]]
KNOWN_FLIP_INIT = [-1, 2, 1, 4, 3, 6, 5, -1]
KNOWN_FLIP_RESULT = [3, 2, 1, 0, 5, 4, 7, 6]


def test_reference_known_flip():
    ro, ci = tg.csr_from_adj(KNOWN_FLIP_ADJ)
    mate, st = oracle.match_forest(ro, ci, np.array(KNOWN_FLIP_INIT, dtype=np.int32))
    assert list(mate) == KNOWN_FLIP_RESULT
    assert st.augmentations == 1
