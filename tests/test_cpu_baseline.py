"""Pins the multi-threaded CPU restatements timed as the CPU baseline
(oracle/msbfs_cpu_baseline.c) before bench.py times them.

Both must produce exactly the tables and traversed-edge counts of the
per-source oracle (itself pinned by networkx fixtures and closed forms in
test_oracle.py) and of the committed networkx fixtures: on every golden graph
(directed ones through the transpose), on graphs whose levels go push-only,
pull-heavy and mixed, for thread counts 1-8 and source counts that are not
multiples of 64, with duplicate sources.
"""
import numpy as np
import pytest

import graphs as tg
import oracle


def _check(ro, ci, src, directed=False, threads=(1, 3, 8)):
    lr, pr, er = oracle.msbfs(ro, ci, src, threads=4)
    tr = oracle.transpose(ro, ci) if directed else None
    for mode in ("dir", "bitpar"):
        for t in threads:
            lv, pa, e = oracle.msbfs_cpu(ro, ci, src, threads=t, mode=mode, transposed=tr)
            assert np.array_equal(lv, lr), f"{mode} threads={t}: levels"
            assert np.array_equal(pa, pr), f"{mode} threads={t}: parents"
            assert e == er, f"{mode} threads={t}: traversed edges {e} != {er}"


@pytest.mark.parametrize("case", list(tg.golden_cases()), ids=lambda c: c[0])
def test_golden_fixtures(case):
    name, ro, ci, src, lv, pa, directed = case
    tr = oracle.transpose(ro, ci) if directed else None
    for mode in ("dir", "bitpar"):
        got_l, got_p, _ = oracle.msbfs_cpu(ro, ci, src, threads=4, mode=mode, transposed=tr)
        assert np.array_equal(got_l, lv) and np.array_equal(got_p, pa), f"{name} {mode}"


def test_closed_forms():
    ro, ci = tg.grid(23, 31)
    _check(ro, ci, np.array([0, 700, 352, 352, 5], dtype=np.int32))
    ro, ci = tg.path(500)  # deep: push-only levels
    _check(ro, ci, np.array([0, 499, 250], dtype=np.int32))
    ro, ci = tg.complete(40)  # one level over nearly all edges: pull
    _check(ro, ci, np.arange(0, 40, 3, dtype=np.int32))
    ro, ci = tg.two_components(30, 40)  # unreached vertices
    _check(ro, ci, np.array([0, 35, 69], dtype=np.int32))


@pytest.mark.parametrize("seed", range(4))
def test_random_graphs_many_sources(seed):
    rng = np.random.default_rng(seed)
    n = 3000 + 1000 * seed
    ro, ci = tg.gnp(n, (2.0 + 3 * seed) / n, 60 + seed)
    src = rng.integers(0, n, size=64 * (seed + 1) + 17).astype(np.int32)  # several words, a partial one
    _check(ro, ci, src, threads=(1, 8))


def test_directed_scale_free():
    for name, ro, ci, src, _, _, directed in tg.golden_cases():
        if directed:
            _check(ro, ci, src, directed=True, threads=(2, 7))


def test_transpose():
    ro, ci = tg.csr_from_adj([[1, 2], [2], [0], []])  # as a directed graph
    io, ic = oracle.transpose(ro, ci)
    assert list(io) == [0, 1, 2, 4, 4] and list(ic) == [2, 0, 0, 1]
