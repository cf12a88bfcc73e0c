"""Synthetic input generators (host-only): determinism, CSR form, and pinned
(n, m, checksum) for the small named workloads."""
import numpy as np
import pytest

import xbmsbfs as xb


def _csr_ok(g: xb.Csr, symmetric: bool):
    ro, ci = g.row_offsets, g.col_indices
    assert ro[0] == 0 and ro[-1] == ci.size and np.all(np.diff(ro) >= 0)
    assert ci.size == 0 or (ci.min() >= 0 and ci.max() < g.n)
    src = np.repeat(np.arange(g.n), np.diff(ro))
    assert not np.any(src == ci), "self-loop"
    # ascending and duplicate-free rows
    same_row = src[1:] == src[:-1]
    assert np.all(ci[1:][same_row] > ci[:-1][same_row])
    if symmetric:
        fwd = src.astype(np.int64) * g.n + ci
        rev = ci.astype(np.int64) * g.n + src
        assert np.array_equal(np.sort(fwd), np.sort(rev))


def test_rmat_small_deterministic():
    a = xb.gen_rmat(10, 8, 0.57, 0.19, 0.19, seed=1, threads=1)
    b = xb.gen_rmat(10, 8, 0.57, 0.19, 0.19, seed=1, threads=4)
    c = xb.gen_rmat(10, 8, 0.57, 0.19, 0.19, seed=2)
    assert a.checksum() == b.checksum() != c.checksum()
    _csr_ok(a, True)


def test_rmat_directed():
    g = xb.gen_rmat(10, 8, 0.6, 0.15, 0.15, seed=7, symmetric=False)
    _csr_ok(g, False)


def test_chung_lu_shape():
    g = xb.gen_chung_lu(20_000, 2.3, 8.4, 400, seed=3)
    _csr_ok(g, True)
    d = g.degrees()
    assert 7.0 < d.mean() < 9.0 and 250 < d.max() < 500


def test_rmat_scale16_pinned():
    """Plain-method workload: R-MAT scale 16, edge factor 16, seed 1; sources seed 2."""
    g, src = xb.workload_rmat_s16()
    assert (g.n, g.m) == (65536, 1819206)
    assert g.checksum() == 0x2A2963EE1ED10E18
    assert src.size == 64 and np.all(g.degrees()[src] > 0) and np.unique(src).size == 64


def test_samplers():
    g = xb.gen_rmat(12, 8, 0.57, 0.19, 0.19, seed=3)
    d = g.degrees()
    u = xb.sample_sources(g, 200, seed=5)
    w = xb.sample_sources(g, 200, seed=5, degree_weighted=True)
    assert np.unique(u).size == 200 and np.unique(w).size == 200
    assert np.all(d[u] > 0) and np.all(d[w] > 0)
    assert d[w].mean() > d[u].mean()
    assert np.array_equal(u, xb.sample_sources(g, 200, seed=5))
