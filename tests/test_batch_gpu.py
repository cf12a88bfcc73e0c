"""Multi-graph batches: one persistent, hipGraph-captured launch (xb_batch_*).

Every graph of a batch must come out bit-exact against the oracle (small
graphs) or against the single-graph engine plus oracle rows (large ones), on
the first run and on every replay of the captured graph.
"""
import numpy as np
import pytest

import graphs as tg
import oracle
import xbmsbfs as xb

pytestmark = pytest.mark.gpu


def _mixed_graphs():
    out = []
    out.append((xb.gen_rmat(8, 8, 0.57, 0.19, 0.19, seed=1), False))
    out.append((xb.gen_rmat(11, 16, 0.57, 0.19, 0.19, seed=2), False))
    out.append((xb.gen_rmat(13, 8, 0.45, 0.25, 0.15, seed=3), False))
    out.append((xb.gen_chung_lu(20_000, 2.1, 12.0, 2_000, seed=4), False))
    out.append((xb.gen_rmat(12, 8, 0.6, 0.15, 0.15, seed=5, symmetric=False), True))
    ro, ci = tg.path(320)  # deeper than the 8-bit level code: plain re-run of its tile
    out.append((xb.Csr(ro, ci), False))
    return out


@pytest.fixture(scope="module")
def mixed():
    rng = np.random.default_rng(7)
    gs = _mixed_graphs()
    hgs = [xb.HGraph(g.row_offsets, g.col_indices, directed=d) for g, d in gs]
    srcs = []
    for g, _ in gs:
        S = int(rng.integers(1, 300))
        s = xb.sample_sources(g, S, seed=int(rng.integers(1 << 30)))
        srcs.append(s if s.size else np.zeros(1, dtype=np.int32))
    srcs[-1] = np.array([0, 319, 160], dtype=np.int32)
    refs = [oracle.msbfs(g.row_offsets, g.col_indices, s, threads=8)[:2] for (g, _), s in zip(gs, srcs)]
    return gs, hgs, srcs, refs


@pytest.mark.parametrize("T", [1, 8, 16])
def test_batch_mixed_vs_oracle(mixed, T):
    gs, hgs, srcs, refs = mixed
    res, st = xb.run_batched(hgs, srcs, words_per_tile=T, pull_cut_small=4, pull_cut_medium=16, pull_cut_hub=128,
                             hub_chunk=64)
    for i, ((lv, pa), (lr, pr)) in enumerate(zip(res, refs)):
        assert np.array_equal(lv, lr), f"graph {i}: levels"
        assert np.array_equal(pa, pr), f"graph {i}: parents"
    assert st.fallback_sources == 3
    assert st.kernel_launches[11] == 1


@pytest.mark.parametrize("fixed_teams", [False, True])
def test_batch_replay_and_single_graph_agree(mixed, fixed_teams):
    """Replays of one batch give the oracle's tables every time, whether the
    grid is re-split between the teams after each run (the default) or not."""
    gs, hgs, srcs, refs = mixed
    lv = [xb.DeviceBuffer((s.size, g.nnodes), fill=7) for s, g in zip(srcs, hgs)]
    pa = [xb.DeviceBuffer((s.size, g.nnodes), fill=7) for s, g in zip(srcs, hgs)]
    b = xb.Batch(hgs, srcs, [t.ptr for t in lv], [t.ptr for t in pa], words_per_tile=4, fixed_teams=fixed_teams)
    for rep in range(5):
        for t in lv + pa:
            t.upload(np.full(t.count, 7, dtype=np.int32))
        st = b.run()
        for i, (lr, pr) in enumerate(refs):
            assert np.array_equal(lv[i].numpy(), lr), f"replay {rep} graph {i}"
            assert np.array_equal(pa[i].numpy(), pr), f"replay {rep} graph {i}"
        assert st.ms_total > 0
    b.close()
    # The single-graph engine on the same inputs.
    for i, (h, s) in enumerate(zip(hgs, srcs)):
        l1, p1 = xb.XBlossomGpuEngine10(h).MultiSourceBFS(s, mode=xb.XB_MODE_BATCHED)
        assert np.array_equal(l1, refs[i][0]) and np.array_equal(p1, refs[i][1])


def test_batch_graph_without_sources(mixed):
    """A graph of a batch may have no sources: it gets no tile and no team,
    and the other graphs' tables stay the oracle's.  A batch with no source
    at all is refused."""
    gs, hgs, srcs, refs = mixed
    srcs2 = list(srcs)
    srcs2[2] = np.zeros(0, dtype=np.int32)
    res, st = xb.run_batched(hgs, srcs2, words_per_tile=4)
    assert st.kernel_launches[11] == 1
    for i, ((lv, pa), (lr, pr)) in enumerate(zip(res, refs)):
        if i == 2:
            assert lv.shape == (0, hgs[2].nnodes) and pa.shape == (0, hgs[2].nnodes)
            continue
        assert np.array_equal(lv, lr), f"graph {i}: levels"
        assert np.array_equal(pa, pr), f"graph {i}: parents"
    with pytest.raises(xb.XbError) as ei:
        xb.run_batched(hgs[:2], [np.zeros(0, dtype=np.int32)] * 2)
    assert ei.value.code == xb.XB_ERR_INVALID_ARG


def test_batch_rejects_bad_sources(mixed):
    gs, hgs, srcs, refs = mixed
    bad = [s.copy() for s in srcs]
    bad[1][0] = hgs[1].nnodes
    with pytest.raises(xb.XbError) as ei:
        xb.run_batched(hgs, bad)
    assert ei.value.code == xb.XB_ERR_SOURCE_RANGE


@pytest.mark.timeout(600)
def test_uneven_batch_full_size():
    """32 R-MAT graphs (scales 12-22, edge factor 8-32, 1-512 sources each) in one
    persistent launch: every graph's tables equal the single-graph engine's
    (device checksums over all S x n entries), and sampled rows equal the oracle."""
    work = xb.workload_uneven_batch()
    hgs = [xb.HGraph(g.row_offsets, g.col_indices) for g, _ in work]
    srcs = [s for _, s in work]
    lv = [xb.DeviceBuffer((s.size, g.nnodes)) for s, g in zip(srcs, hgs)]
    pa = [xb.DeviceBuffer((s.size, g.nnodes)) for s, g in zip(srcs, hgs)]
    b = xb.Batch(hgs, srcs, [t.ptr for t in lv], [t.ptr for t in pa])
    st = b.run()
    assert st.kernel_launches[11] == 1 and st.tiles >= len(work)
    for i, ((g, s), h) in enumerate(zip(work, hgs)):
        eng = xb.XBlossomGpuEngine10(h)
        l1 = xb.DeviceBuffer((s.size, g.n))
        p1 = xb.DeviceBuffer((s.size, g.n))
        eng.MultiSourceBFS(s, l1.ptr, p1.ptr, mode=xb.XB_MODE_BATCHED, device_out=True)
        assert lv[i].checksum() == l1.checksum(), f"graph {i}: levels"
        assert pa[i].checksum() == p1.checksum(), f"graph {i}: parents"
        l1.free()
        p1.free()
        rows = sorted({0, s.size - 1, s.size // 2})
        lr, pr, _ = oracle.msbfs(g.row_offsets, g.col_indices, s[rows], threads=16)
        got_l = np.concatenate([lv[i].rows(r, r + 1) for r in rows])
        got_p = np.concatenate([pa[i].rows(r, r + 1) for r in rows])
        assert np.array_equal(got_l, lr) and np.array_equal(got_p, pr), f"graph {i}: oracle rows"
    # Later replays run on grid splits re-balanced from the team clocks: the
    # tables stay the same.
    sums = [(t.checksum(), u.checksum()) for t, u in zip(lv, pa)]
    for _ in range(3):
        b.run()
    assert sums == [(t.checksum(), u.checksum()) for t, u in zip(lv, pa)]
    b.close()


def test_batch_refuses_run_after_layout_rebuild(mixed):
    """A batch holds its graphs' pull-layout arrays.  A run with other
    cut-offs rebuilds (frees and reallocates) them; a run with the original
    cut-offs afterwards rebuilds them again.  Equal cut values then do not mean
    the same arrays, so the batch refuses to run before launching anything,
    and a new batch runs bit-exact."""
    gs, hgs, srcs, refs = mixed
    i = 1
    lv = xb.DeviceBuffer((srcs[i].size, hgs[i].nnodes))
    pa = xb.DeviceBuffer((srcs[i].size, hgs[i].nnodes))
    b = xb.Batch([hgs[i]], [srcs[i]], [lv.ptr], [pa.ptr], words_per_tile=4)
    b.run()
    assert np.array_equal(pa.numpy(), refs[i][1])
    eng = xb.XBlossomGpuEngine10(hgs[i])
    eng.MultiSourceBFS(srcs[i][:5], mode=xb.XB_MODE_BATCHED, pull_cut_small=4, pull_cut_medium=16)
    eng.MultiSourceBFS(srcs[i][:5], mode=xb.XB_MODE_BATCHED)  # default cut-offs again
    with pytest.raises(xb.XbError) as ei:
        b.run()
    assert ei.value.code == xb.XB_ERR_INVALID_ARG
    b.close()
    b2 = xb.Batch([hgs[i]], [srcs[i]], [lv.ptr], [pa.ptr], words_per_tile=4)
    lv.upload(np.full(lv.count, 7, dtype=np.int32))
    b2.run()
    assert np.array_equal(lv.numpy(), refs[i][0]) and np.array_equal(pa.numpy(), refs[i][1])
    b2.close()


@pytest.mark.parametrize("emit_inside", [False, True])
def test_batch_emit_after_kernel_and_inside(emit_inside):
    """A graph whose sources fit one tile is emitted by standalone launches
    captured after the persistent kernel, gated on the depth the kernel
    leaves in the job table (4 planes up to level 14, 8 beyond); with
    XB_BATCH_EMIT_INSIDE every team emits inside the kernel.  Both give the
    oracle's tables on every replay: a 99-level path (the 8-plane launch), a
    symmetric and a directed R-MAT, and a graph of two tiles (emitted inside
    the kernel either way)."""
    ro, ci = tg.path(100)
    gs = [(xb.Csr(ro, ci), False),
          (xb.gen_rmat(11, 16, 0.57, 0.19, 0.19, seed=12), False),
          (xb.gen_rmat(12, 8, 0.6, 0.15, 0.15, seed=13, symmetric=False), True),
          (xb.gen_chung_lu(20_000, 2.1, 12.0, 2_000, seed=14), False)]
    counts = [3, 100, 40, 300]  # 300 sources at 4 words per tile: two tiles
    hgs = [xb.HGraph(g.row_offsets, g.col_indices, directed=d) for g, d in gs]
    srcs = [np.array([0, 50, 99], dtype=np.int32)]
    srcs += [xb.sample_sources(g, c, seed=20 + i) for i, ((g, _), c) in enumerate(zip(gs[1:], counts[1:]))]
    refs = [oracle.msbfs(g.row_offsets, g.col_indices, s, threads=8)[:2] for (g, _), s in zip(gs, srcs)]
    lv = [xb.DeviceBuffer((s.size, g.nnodes)) for s, g in zip(srcs, hgs)]
    pa = [xb.DeviceBuffer((s.size, g.nnodes)) for s, g in zip(srcs, hgs)]
    b = xb.Batch(hgs, srcs, [t.ptr for t in lv], [t.ptr for t in pa], words_per_tile=4, emit_inside=emit_inside)
    for rep in range(3):
        for t in lv + pa:
            t.upload(np.full(t.count, 7, dtype=np.int32))
        st = b.run()
        assert st.kernel_launches[11] == 1 and st.max_level == 99
        for i, (lr, pr) in enumerate(refs):
            assert np.array_equal(lv[i].numpy(), lr), f"replay {rep} graph {i}: levels"
            assert np.array_equal(pa[i].numpy(), pr), f"replay {rep} graph {i}: parents"
    b.close()
    for t in lv + pa:
        t.free()


def test_batch_emit_gate_at_the_plane_boundary():
    """The captured emits after the persistent kernel split at the depth the
    kernel records: 4 planes up to level 14, 8 from level 15.  Paths whose
    tiles end exactly at levels 13, 14, 15 and 16 (one graph each, one tile
    each) in one batch come out equal to the oracle, on two replays."""
    gs = [xb.Csr(*tg.path(k)) for k in (14, 15, 16, 17)]
    hgs = [xb.HGraph(g.row_offsets, g.col_indices) for g in gs]
    srcs = [np.array([0, g.n // 2], dtype=np.int32) for g in gs]
    refs = [oracle.msbfs(g.row_offsets, g.col_indices, s, threads=2)[:2] for g, s in zip(gs, srcs)]
    assert [int(r[0].max()) for r in refs] == [13, 14, 15, 16]
    lv = [xb.DeviceBuffer((s.size, g.nnodes)) for s, g in zip(srcs, hgs)]
    pa = [xb.DeviceBuffer((s.size, g.nnodes)) for s, g in zip(srcs, hgs)]
    b = xb.Batch(hgs, srcs, [t.ptr for t in lv], [t.ptr for t in pa], words_per_tile=1)
    for rep in range(2):
        for t in lv + pa:
            t.upload(np.full(t.count, 7, dtype=np.int32))
        st = b.run()
        assert st.max_level == 16
        for i, (lr, pr) in enumerate(refs):
            assert np.array_equal(lv[i].numpy(), lr), f"replay {rep} path {i}: levels"
            assert np.array_equal(pa[i].numpy(), pr), f"replay {rep} path {i}: parents"
    b.close()
    for t in lv + pa:
        t.free()
