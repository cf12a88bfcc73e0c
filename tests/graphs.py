"""Test graphs with closed-form BFS levels and lowest-id parents, plus an
independent invariant checker.

The closed forms below are written from the graph structure (distances on a
path, cycle, grid, tree, ...), not from the oracle's algorithm, so they pin the
oracle and the engine independently.
"""
from __future__ import annotations

import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def csr_from_adj(adj):
    n = len(adj)
    ro = np.zeros(n + 1, dtype=np.int32)
    cols = []
    for v in range(n):
        nb = sorted(set(adj[v]))
        cols.extend(nb)
        ro[v + 1] = ro[v] + len(nb)
    return ro, np.array(cols, dtype=np.int32)


# ---- graphs -------------------------------------------------------------

def path(n):
    return csr_from_adj([[u for u in (v - 1, v + 1) if 0 <= u < n] for v in range(n)])


def cycle(n):
    return csr_from_adj([[(v - 1) % n, (v + 1) % n] for v in range(n)])


def star(n):
    return csr_from_adj([list(range(1, n))] + [[0] for _ in range(1, n)])


def complete(n):
    return csr_from_adj([[u for u in range(n) if u != v] for v in range(n)])


def grid(R, C):
    adj = []
    for r in range(R):
        for c in range(C):
            nb = []
            if r > 0:
                nb.append((r - 1) * C + c)
            if r < R - 1:
                nb.append((r + 1) * C + c)
            if c > 0:
                nb.append(r * C + c - 1)
            if c < C - 1:
                nb.append(r * C + c + 1)
            adj.append(nb)
    return csr_from_adj(adj)


def binary_tree(n):
    adj = [[] for _ in range(n)]
    for i in range(1, n):
        p = (i - 1) // 2
        adj[i].append(p)
        adj[p].append(i)
    return csr_from_adj(adj)


def two_components(a, b):
    """Path of a vertices, then a disjoint path of b vertices."""
    adj = [[u for u in (v - 1, v + 1) if 0 <= u < a] for v in range(a)]
    adj += [[a + u for u in (v - 1, v + 1) if 0 <= u < b] for v in range(b)]
    return csr_from_adj(adj)


def complete_bipartite(a, b):
    """K_{a,b}: left 0..a-1, right a..a+b-1."""
    return csr_from_adj([list(range(a, a + b)) for _ in range(a)] + [list(range(a)) for _ in range(b)])


def petersen():
    outer = [[(i + 1) % 5, (i - 1) % 5, i + 5] for i in range(5)]
    inner = [[5 + (i + 2) % 5, 5 + (i - 2) % 5, i] for i in range(5)]
    return csr_from_adj(outer + inner)


def from_edges(n, u, v):
    """Symmetric CSR (ascending, deduplicated, no self-loops) from an edge list."""
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    keep = u != v
    a = np.concatenate([u[keep], v[keep]])
    b = np.concatenate([v[keep], u[keep]])
    key = np.unique(a * n + b)
    src, dst = key // n, key % n
    ro = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ro, src + 1, 1)
    return np.cumsum(ro).astype(np.int32), dst.astype(np.int32)


def gnp(n, p, seed):
    rng = np.random.default_rng(seed)
    m = rng.binomial(n * (n - 1) // 2, p)
    return from_edges(n, rng.integers(0, n, m), rng.integers(0, n, m))


def random_bipartite(nl, nr, avg_deg, seed):
    """Left 0..nl-1, right nl..nl+nr-1, about avg_deg edges per left vertex."""
    rng = np.random.default_rng(seed)
    m = int(nl * avg_deg)
    return from_edges(nl + nr, rng.integers(0, nl, m), nl + rng.integers(0, nr, m))


def nested_odd_cycles(k, depth, seed):
    """Odd cycles of odd cycles (k parts per level, `depth` levels) joined at
    random vertices, plus pendant vertices and random chords: blossoms nested
    inside blossoms."""
    rng = np.random.default_rng(seed)
    edges = []
    count = [0]

    def build(d):
        if d == 0:
            count[0] += 1
            return [count[0] - 1]
        parts = [build(d - 1) for _ in range(k)]
        for i in range(k):
            a, b = parts[i], parts[(i + 1) % k]
            edges.append((a[rng.integers(len(a))], b[rng.integers(len(b))]))
        return [v for part in parts for v in part]

    build(depth)
    n = count[0]
    for _ in range(n // 4):
        edges.append((int(rng.integers(n)), count[0]))
        count[0] += 1
    n = count[0]
    for _ in range(n // 5):
        edges.append((int(rng.integers(n)), int(rng.integers(n))))
    e = np.array(edges, dtype=np.int64)
    return from_edges(n, e[:, 0], e[:, 1])


def with_loops_and_duplicates(ro, ci, seed):
    """The same graph with self-loops and repeated entries added (raw CSR input
    may hold both); lists stay sorted."""
    rng = np.random.default_rng(seed)
    n = ro.size - 1
    adj = [list(ci[ro[v]:ro[v + 1]]) for v in range(n)]
    for v in rng.choice(n, size=max(1, n // 5), replace=False):
        adj[v].append(v)
    for v in range(n):
        if adj[v] and rng.random() < 0.3:
            adj[v].append(adj[v][int(rng.integers(len(adj[v])))])
    ro2 = np.zeros(n + 1, dtype=np.int32)
    cols = []
    for v in range(n):
        a = sorted(int(x) for x in adj[v])
        cols.extend(a)
        ro2[v + 1] = ro2[v] + len(a)
    return ro2, np.array(cols, dtype=np.int32)


def matching_size(mate):
    mate = np.asarray(mate)
    return int(np.count_nonzero(mate >= 0)) // 2


def check_matching(ro, ci, mate):
    """Asserts the matching rules (ValidateMatchingGPU, apps/xb/test/include/
    common/test_utils.h:8-19, with -1 for exposed): symmetric, in range, and
    every pair an edge of the graph."""
    mate = np.asarray(mate)
    n = ro.size - 1
    assert mate.shape == (n,)
    assert np.all(mate >= -1) and np.all(mate < n)
    m = np.flatnonzero(mate >= 0)
    assert np.array_equal(mate[mate[m]], m), "matching not symmetric"
    assert not np.any(mate[m] == m), "vertex matched to itself"
    if m.size:
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(ro))
        keys = rows * n + ci.astype(np.int64)  # ascending: rows ascending, lists ascending
        q = m.astype(np.int64) * n + mate[m]
        k = np.minimum(np.searchsorted(keys, q), keys.size - 1)
        bad = m[keys[k] != q]
        assert bad.size == 0, f"pair ({bad[0]}, {mate[bad[0]]}) is not an edge"


# ---- closed forms (level, parent) ---------------------------------------

def expect_path(n, s):
    v = np.arange(n)
    lv = np.abs(v - s)
    pa = np.where(v > s, v - 1, v + 1)
    pa[s] = s
    return lv, pa


def expect_cycle(n, s):
    v = np.arange(n)
    fwd = (v - s) % n  # steps going up from s
    lv = np.minimum(fwd, n - fwd)
    pa = np.empty(n, dtype=np.int64)
    for x in range(n):
        if x == s:
            pa[x] = s
            continue
        cands = [u for u in ((x - 1) % n, (x + 1) % n) if lv[u] == lv[x] - 1]
        pa[x] = min(cands)
    return lv, pa


def expect_star(n, s):
    lv = np.full(n, 2 if s != 0 else 1)
    pa = np.zeros(n, dtype=np.int64)
    if s == 0:
        lv[0] = 0
    else:
        lv[0] = 1
        lv[s] = 0
        pa[0] = s
        pa[s] = s
    return lv, pa


def expect_complete(n, s):
    lv = np.ones(n, dtype=np.int64)
    lv[s] = 0
    pa = np.full(n, s)
    return lv, pa


def expect_grid(R, C, s):
    rs, cs = divmod(s, C)
    lv = np.empty(R * C, dtype=np.int64)
    pa = np.empty(R * C, dtype=np.int64)
    for v in range(R * C):
        r, c = divmod(v, C)
        lv[v] = abs(r - rs) + abs(c - cs)
        cands = []
        if r > rs:
            cands.append(v - C)
        if r < rs:
            cands.append(v + C)
        if c > cs:
            cands.append(v - 1)
        if c < cs:
            cands.append(v + 1)
        pa[v] = min(cands) if cands else s
    return lv, pa


def expect_binary_tree(n, s):
    def anc(x):
        out = [x]
        while x:
            x = (x - 1) // 2
            out.append(x)
        return out

    As = anc(s)
    lv = np.empty(n, dtype=np.int64)
    pa = np.empty(n, dtype=np.int64)
    for v in range(n):
        Av = anc(v)
        common = next(x for x in Av if x in set(As))
        lv[v] = Av.index(common) + As.index(common)
        if v == s:
            pa[v] = s
        elif common == v:
            # v is an ancestor of s: step back down towards s
            pa[v] = As[As.index(v) - 1]
        else:
            pa[v] = (v - 1) // 2
    return lv, pa


CLOSED_FORMS = {
    "path97": (lambda: path(97), lambda s: expect_path(97, s), [0, 48, 96, 13]),
    "cycle40": (lambda: cycle(40), lambda s: expect_cycle(40, s), [0, 7, 39, 20]),
    "cycle41": (lambda: cycle(41), lambda s: expect_cycle(41, s), [0, 17, 40]),
    "star30": (lambda: star(30), lambda s: expect_star(30, s), [0, 1, 29, 15]),
    "complete25": (lambda: complete(25), lambda s: expect_complete(25, s), [0, 12, 24]),
    "grid9x13": (lambda: grid(9, 13), lambda s: expect_grid(9, 13, s), [0, 58, 116, 12, 104, 30]),
    "bintree127": (lambda: binary_tree(127), lambda s: expect_binary_tree(127, s), [0, 1, 2, 63, 100, 126]),
}


def closed_form_cases():
    for name, (mk, ex, srcs) in CLOSED_FORMS.items():
        ro, ci = mk()
        n = ro.size - 1
        lv = np.stack([ex(s)[0] for s in srcs]).astype(np.int32)
        pa = np.stack([ex(s)[1] for s in srcs]).astype(np.int32)
        yield name, ro, ci, np.array(srcs, dtype=np.int32), lv, pa


def golden_cases():
    for path_ in sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))):
        z = np.load(path_)  # allow_pickle defaults to False
        yield (os.path.basename(path_)[:-4], z["row_offsets"], z["col_indices"], z["sources"], z["levels"],
               z["parents"], bool(z["directed"][0]))


# ---- invariant checker ---------------------------------------------------

def check_invariants(ro, ci, sources, levels, parents, rows=None):
    """Asserts the definition of BFS levels + lowest-id parents row by row.

    For each checked source row: level[src] = 0, parent[src] = src; every edge
    (u->v) out of a reached u reaches v with level[v] <= level[u] + 1; every
    reached v != src has level >= 1 and parent = the minimum in-neighbour u
    with level[u] == level[v] - 1; unreached vertices have parent -1.
    Vectorised over edges, so it runs at full graph size for a few rows.
    """
    n = ro.size - 1
    deg = np.diff(ro)
    src_of_edge = np.repeat(np.arange(n, dtype=np.int64), deg)
    dst = ci.astype(np.int64)
    rows = range(len(sources)) if rows is None else rows
    for r in rows:
        s = int(sources[r])
        L = levels[r].astype(np.int64)
        P = parents[r].astype(np.int64)
        assert L[s] == 0 and P[s] == s, f"row {r}: source entry wrong"
        reached = L >= 0
        assert np.all(P[~reached] == -1), f"row {r}: unreached vertex has a parent"
        ru = reached[src_of_edge]
        assert np.all(reached[dst[ru]]), f"row {r}: edge from a reached vertex to an unreached one"
        assert np.all(L[dst[ru]] <= L[src_of_edge[ru]] + 1), f"row {r}: level jumps over an edge"
        others = reached.copy()
        others[s] = False
        assert np.all(L[others] >= 1), f"row {r}: a non-source vertex at level 0"
        # lowest-id in-neighbour one level closer
        good = ru & (L[dst] == L[src_of_edge] + 1) & (dst != s)
        best = np.full(n, np.iinfo(np.int64).max)
        np.minimum.at(best, dst[good], src_of_edge[good])
        best[s] = s
        assert np.all(P[others] == best[others]), f"row {r}: parent is not the lowest-id predecessor"
