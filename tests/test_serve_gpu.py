"""Serving: continuous batching over one device-resident graph (xb_server_*).

Each request's level and parent tables must be the same whether the request
ran alone or packed into tiles with others (device checksums over all S x n
entries against the single-graph engine), and bit-exact against the oracle
for sampled requests.  Admission must hold the budget, and the error paths
must report the documented status codes.
"""
import ctypes

import numpy as np
import pytest

import graphs as tg
import oracle
import xbmsbfs as xb

pytestmark = pytest.mark.gpu


def _requests(g, count, seed, smax=256):
    rng = np.random.default_rng(seed)
    deg = g.degrees()
    cand = np.flatnonzero(deg > 0).astype(np.int32)
    out = []
    for _ in range(count):
        S = int(np.floor(np.exp(rng.uniform(0.0, np.log(smax + 1.0)))))
        S = max(1, min(smax, S))
        out.append(rng.choice(cand, size=S, replace=True).astype(np.int32))
    return out


def _host_tables(info):
    cells = info.S * info.n
    lv = np.empty(cells, dtype=np.int32)
    pa = np.empty(cells, dtype=np.int32)
    lib = xb.load_engine()
    xb._check(lib.xb_memcpy_d2h(lv.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(info.levels), cells * 4))
    xb._check(lib.xb_memcpy_d2h(pa.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(info.parents), cells * 4))
    return lv.reshape(info.S, info.n), pa.reshape(info.S, info.n)


@pytest.fixture(scope="module")
def web():
    g = xb.gen_rmat_capped(15, 30_000, 400_000, 0.6, 0.15, 0.15, seed=7)
    return g, xb.HGraph(g.row_offsets, g.col_indices, directed=True)


def _alone_checksums(hg, src, **tunables):
    n = hg.nnodes
    lv = xb.DeviceBuffer((src.size, n))
    pa = xb.DeviceBuffer((src.size, n))
    xb.XBlossomGpuEngine10(hg).MultiSourceBFS(src, lv.ptr, pa.ptr, mode=xb.XB_MODE_BATCHED, device_out=True,
                                              **tunables)
    out = (lv.checksum(), pa.checksum())
    lv.free()
    pa.free()
    return out


@pytest.mark.parametrize("T", [1, 8, 16])
def test_mixed_equals_alone(web, T):
    g, hg = web
    reqs = _requests(g, 40, seed=11 + T)
    srv = xb.Server(hg, slots=3, budget_bytes=4 << 30, words_per_tile=T)
    ids = [srv.submit(s) for s in reqs]
    mixed = []
    for i, rid in enumerate(ids):
        info = srv.wait(rid, timeout_ms=60_000)
        assert info.status == xb.XB_OK and info.S == reqs[i].size and info.n == g.n
        cells = info.S * info.n
        mixed.append((xb.device_checksum(info.levels, cells), xb.device_checksum(info.parents, cells),
                      info.edges, info.tiles))
        if i in (0, 7, 23):
            lv, pa = _host_tables(info)
            lr, pr, edges = oracle.msbfs(g.row_offsets, g.col_indices, reqs[i], threads=8)
            assert np.array_equal(lv, lr), f"request {i}: levels"
            assert np.array_equal(pa, pr), f"request {i}: parents"
            assert info.edges == edges, f"request {i}: traversed edges"
        srv.release(rid)
    c = srv.counters()
    assert c.completed == len(reqs) and c.released == len(reqs) and c.failed == 0
    assert c.used_bytes == 0 and c.queued == 0
    # Rows of several requests shared tiles.
    assert c.tiles < len(reqs) or T == 1
    srv.close()
    for i, s in enumerate(reqs):
        assert _alone_checksums(hg, s, words_per_tile=T) == mixed[i][:2], f"request {i}: mixed != alone"


def test_request_larger_than_a_tile(web):
    g, hg = web
    cand = np.flatnonzero(g.degrees() > 0).astype(np.int32)
    src = np.random.default_rng(3).choice(cand, size=200, replace=True).astype(np.int32)
    srv = xb.Server(hg, slots=2, budget_bytes=1 << 30, words_per_tile=1)  # 64 rows per tile
    rid = srv.submit(src)
    info = srv.wait(rid, timeout_ms=60_000)
    assert info.S == 200 and info.tiles == 4
    lv, pa = _host_tables(info)
    lr, pr, edges = oracle.msbfs(g.row_offsets, g.col_indices, src, threads=8)
    assert np.array_equal(lv, lr) and np.array_equal(pa, pr) and info.edges == edges
    srv.release(rid)
    srv.close()


def test_deep_tile_takes_the_plain_path():
    ro, ci = tg.path(400)
    hg = xb.HGraph(ro, ci)
    srv = xb.Server(hg, slots=2, budget_bytes=64 << 20)
    reqs = [np.array([0, 399], dtype=np.int32), np.array([200], dtype=np.int32), np.array([5, 6, 7], dtype=np.int32)]
    ids = [srv.submit(s) for s in reqs]
    for s, rid in zip(reqs, ids):
        info = srv.wait(rid, timeout_ms=60_000)
        lv, pa = _host_tables(info)
        lr, pr, _ = oracle.msbfs(ro, ci, s)
        assert np.array_equal(lv, lr) and np.array_equal(pa, pr)
        srv.release(rid)
    srv.close()


def test_admission_holds_the_budget(web):
    g, hg = web
    reqs = _requests(g, 12, seed=5, smax=64)
    biggest = max(8 * s.size * g.n for s in reqs)
    budget = 2 * biggest + 4096
    srv = xb.Server(hg, slots=3, budget_bytes=budget)
    ids = [srv.submit(s) for s in reqs]
    for i, rid in enumerate(ids):
        info = srv.wait(rid, timeout_ms=60_000)
        assert info.status == xb.XB_OK
        assert xb.device_checksum(info.parents, info.S * info.n) == _alone_checksums(hg, reqs[i])[1]
        srv.release(rid)
    c = srv.counters()
    assert c.peak_used_bytes <= budget and c.completed == len(reqs)
    srv.close()


def test_waiting_request_times_out_until_memory_frees(web):
    g, hg = web
    a = _requests(g, 1, seed=9, smax=16)[0]
    budget = 8 * a.size * g.n + 256
    srv = xb.Server(hg, slots=2, budget_bytes=budget)
    ra = srv.submit(a)
    srv.wait(ra, timeout_ms=60_000)
    rb = srv.submit(a)  # cannot be admitted while a's tables are held
    with pytest.raises(xb.XbError) as ei:
        srv.wait(rb, timeout_ms=200)
    assert ei.value.code == xb.XB_ERR_TIMEOUT
    srv.release(ra)
    info = srv.wait(rb, timeout_ms=60_000)
    assert info.status == xb.XB_OK
    srv.release(rb)
    srv.close()


def test_rejections(web):
    g, hg = web
    srv = xb.Server(hg, slots=1, budget_bytes=8 * 4 * g.n)
    with pytest.raises(xb.XbError) as ei:
        srv.submit(np.arange(5, dtype=np.int32))  # 5 rows need more than the 4-row budget
    assert ei.value.code == xb.XB_ERR_BUDGET
    with pytest.raises(xb.XbError) as ei:
        srv.submit(np.array([g.n], dtype=np.int32))
    assert ei.value.code == xb.XB_ERR_SOURCE_RANGE
    with pytest.raises(xb.XbError) as ei:
        srv.submit(np.array([], dtype=np.int32))
    assert ei.value.code == xb.XB_ERR_INVALID_ARG
    with pytest.raises(xb.XbError) as ei:
        srv.wait(12345, timeout_ms=10)
    assert ei.value.code == xb.XB_ERR_INVALID_ARG
    srv.close()
    with pytest.raises(xb.XbError) as ei:
        xb.Server(hg, mode=xb.XB_MODE_PLAIN)
    assert ei.value.code == xb.XB_ERR_UNSUPPORTED


def test_server_refuses_tiles_after_layout_rebuild(web):
    """The server's slots read the graph's pull-layout arrays.  Re-tuning the
    graph while the server lives (here while it is idle: other cut-offs, then
    the original ones back) rebuilds them, so later tiles fail with
    XB_ERR_INVALID_ARG instead of reading freed memory."""
    g, hg = web
    reqs = _requests(g, 2, seed=21, smax=8)
    srv = xb.Server(hg, slots=1, budget_bytes=256 << 20)
    rid = srv.submit(reqs[0])
    assert srv.wait(rid, timeout_ms=60_000).status == xb.XB_OK
    srv.release(rid)
    eng = xb.XBlossomGpuEngine10(hg)
    eng.MultiSourceBFS(reqs[1], mode=xb.XB_MODE_BATCHED, pull_cut_small=4, pull_cut_medium=16)
    eng.MultiSourceBFS(reqs[1], mode=xb.XB_MODE_BATCHED)
    rid = srv.submit(reqs[1])
    with pytest.raises(xb.XbError) as ei:
        srv.wait(rid, timeout_ms=60_000)
    assert ei.value.code == xb.XB_ERR_INVALID_ARG
    srv.release(rid)
    srv.close()


@pytest.mark.timeout(900)
def test_web_crawl_full_size():
    """The serving workload at its configured size: a directed R-MAT web-crawl
    graph (n = 20M, ~300M edges, a=.6 b=.15 c=.15, seed 7) resident on the
    device, queries of 1-256 sources (log-uniform, seed 8) arriving by a
    Poisson process at 2,000/s (seed 7).  Every request's tables, computed in
    tiles shared with other requests, equal the tables xb_msbfs_run writes for
    the request alone (device checksums over all S x n entries) and its
    traversed-edge count; sampled rows of three requests equal the oracle, and
    one row satisfies the definition invariants."""
    import time

    g = xb.gen_rmat_capped(25, 20_000_000, 386_000_000, 0.6, 0.15, 0.15, seed=7)
    assert g.n == 20_000_000 and 250_000_000 < g.m < 350_000_000
    hg = xb.HGraph(g.row_offsets, g.col_indices, directed=True)
    rng_a = np.random.default_rng(7)
    rng_s = np.random.default_rng(8)
    cand = np.flatnonzero(g.degrees() > 0).astype(np.int32)
    N = 24
    arrivals = np.cumsum(rng_a.exponential(1.0 / 2000.0, size=N))
    reqs = []
    for _ in range(N):
        S = max(1, min(256, int(np.floor(np.exp(rng_s.uniform(0.0, np.log(257.0)))))))
        reqs.append(rng_s.choice(cand, size=S, replace=True).astype(np.int32))
    srv = xb.Server(hg, slots=2, words_per_tile=8)
    t0 = time.perf_counter()
    ids = []
    for i in range(N):
        delay = t0 + arrivals[i] - arrivals[0] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(reqs[i]))
    mixed = []
    for i, rid in enumerate(ids):
        info = srv.wait(rid, timeout_ms=300_000)
        assert info.status == xb.XB_OK and info.S == reqs[i].size and info.n == g.n
        cells = info.S * info.n
        mixed.append((xb.device_checksum(info.levels, cells), xb.device_checksum(info.parents, cells), info.edges))
        srv.release(rid)
    c = srv.counters()
    assert c.completed == N and c.failed == 0 and c.used_bytes == 0
    assert c.tiles < N  # rows of several requests shared tiles
    srv.close()
    eng = xb.XBlossomGpuEngine10(hg)
    checked = {0: [0], N // 2: [0], N - 1: [0]}
    for i, s in enumerate(reqs):
        lv = xb.DeviceBuffer((s.size, g.n))
        pa = xb.DeviceBuffer((s.size, g.n))
        eng.MultiSourceBFS(s, lv.ptr, pa.ptr, mode=xb.XB_MODE_BATCHED, device_out=True, words_per_tile=8)
        assert (lv.checksum(), pa.checksum()) == mixed[i][:2], f"request {i}: mixed != alone"
        assert eng.CountTraversed(lv.ptr, s.size) == mixed[i][2], f"request {i}: traversed edges"
        if i in checked:
            rows = sorted({0, s.size - 1})
            lr, pr, _ = oracle.msbfs(g.row_offsets, g.col_indices, s[rows], threads=16)
            got_l = np.concatenate([lv.rows(r, r + 1) for r in rows])
            got_p = np.concatenate([pa.rows(r, r + 1) for r in rows])
            assert np.array_equal(got_l, lr) and np.array_equal(got_p, pr), f"request {i}: oracle rows"
            if i == 0:
                tg.check_invariants(g.row_offsets, g.col_indices, s[rows[:1]], got_l[:1], got_p[:1])
        lv.free()
        pa.free()
