"""Writes the networkx golden fixtures for the BFS level / lowest-id parent rule.

Run in a container that has networkx (the committed .npz files are what the
tests read; networkx itself is not needed to run the tests):

    python tests/golden/make_fixtures.py

For every graph and source:
  levels[s][v]  = networkx.single_source_shortest_path_length(G, src)[v], -1 if absent
  parents[s][v] = min(networkx.predecessor(G, src)[v])   (shortest-path predecessors),
                  src for v == src, -1 if unreached.
networkx follows out-edges of a DiGraph, so directed fixtures pin the in-edge
(predecessor) side of the rule.  Graphs are stored as int32 CSR with ascending
adjacency.
"""
from __future__ import annotations

import os

import networkx as nx
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def to_csr(G, n):
    ro = np.zeros(n + 1, dtype=np.int32)
    cols = []
    for v in range(n):
        nb = sorted(set(G.successors(v)) if G.is_directed() else set(G.neighbors(v)))
        cols.extend(nb)
        ro[v + 1] = ro[v] + len(nb)
    return ro, np.array(cols, dtype=np.int32)


def expected(G, n, sources):
    S = len(sources)
    lv = np.full((S, n), -1, dtype=np.int32)
    pa = np.full((S, n), -1, dtype=np.int32)
    for i, s in enumerate(sources):
        dist = nx.single_source_shortest_path_length(G, int(s))
        pred = nx.predecessor(G, int(s))
        for v, d in dist.items():
            lv[i, v] = d
            pa[i, v] = int(s) if v == s else min(pred[v])
    return lv, pa


def write(name, G, n, sources, directed):
    ro, ci = to_csr(G, n)
    lv, pa = expected(G, n, sources)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), row_offsets=ro, col_indices=ci,
                        sources=np.asarray(sources, dtype=np.int32), levels=lv, parents=pa,
                        directed=np.array([1 if directed else 0], dtype=np.int32))
    print(f"{name}: n={n} m={ci.size} S={len(sources)} maxlevel={lv.max()}")


def main():
    rng = np.random.default_rng(20261203)
    # Sparse undirected random graph with isolated vertices and several components.
    G = nx.gnp_random_graph(300, 0.012, seed=1)
    write("gnp300", G, 300, list(rng.choice(300, 70, replace=False)) + [0, 0, 299], False)
    # Denser undirected graph: many equal-level candidates, so the lowest-id choice matters.
    G = nx.gnp_random_graph(200, 0.08, seed=2)
    write("gnp200_dense", G, 200, list(range(0, 200, 3)), False)
    # Directed graphs: parents must come from in-edges.
    G = nx.gnp_random_graph(250, 0.02, seed=3, directed=True)
    write("digraph250", G, 250, list(rng.choice(250, 80, replace=False)), True)
    G = nx.scale_free_graph(400, seed=4)
    D = nx.DiGraph()
    D.add_nodes_from(range(400))
    D.add_edges_from((u, v) for u, v in G.edges() if u != v)
    write("scalefree400_directed", D, 400, list(rng.choice(400, 65, replace=False)), True)
    # Power-law undirected (Barabasi-Albert) with hubs.
    G = nx.barabasi_albert_graph(500, 3, seed=5)
    write("ba500", G, 500, list(rng.choice(500, 130, replace=False)), False)
    # Long thin graph: deep levels (lollipop: clique + path).
    G = nx.lollipop_graph(12, 120)
    write("lollipop", G, 132, [0, 5, 131, 70, 11], False)


if __name__ == "__main__":
    main()
