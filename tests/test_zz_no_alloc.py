"""The allocation-free entry point (xb_msbfs_run_noalloc).

Its tables must equal xb_msbfs_run's for the same sources and options, on
two calls in a row where the second is smaller (the static device array keeps
the first call's contents, which must not leak into the second), in every
engine mode; oversize graphs and unsupported options must be refused with the
documented status codes.
"""
import ctypes

import numpy as np
import pytest

import graphs as tg
import xbmsbfs as xb

pytestmark = pytest.mark.gpu


def _tables(hg, S):
    return xb.DeviceBuffer((S, hg.nnodes), fill=-7), xb.DeviceBuffer((S, hg.nnodes), fill=-7)


def _reference(hg, src, **kw):
    lv, pa = _tables(hg, src.size)
    xb.XBlossomGpuEngine10(hg).MultiSourceBFS(src, lv.ptr, pa.ptr, device_out=True, **kw)
    out = (lv.numpy(), pa.numpy())
    lv.free()
    pa.free()
    return out


@pytest.fixture(scope="module")
def graphs():
    a = xb.gen_rmat(14, 8, 0.57, 0.19, 0.19, seed=31)
    b = xb.gen_chung_lu(150_000, 2.1, 12.0, 4_000, seed=32)
    c = xb.gen_rmat(12, 8, 0.6, 0.15, 0.15, seed=33, symmetric=False)
    return [(a, xb.HGraph(a.row_offsets, a.col_indices)), (b, xb.HGraph(b.row_offsets, b.col_indices)),
            (c, xb.HGraph(c.row_offsets, c.col_indices, directed=True))]


@pytest.mark.parametrize("mode,T", [(xb.XB_MODE_PLAIN, 8), (xb.XB_MODE_BATCHED, 1), (xb.XB_MODE_BATCHED, 4),
                                    (xb.XB_MODE_AUTO, 16)])
def test_noalloc_equals_msbfs_run(graphs, mode, T):
    for gi, (g, hg) in enumerate(graphs):
        # Two calls in a row, the second smaller, on different sources.
        for call, (S, seed) in enumerate([(300, 41 + gi), (37, 51 + gi)]):
            src = xb.sample_sources(g, S, seed=seed)
            lv, pa = _tables(hg, src.size)
            st = xb.msbfs_run_noalloc(hg, src, lv.ptr, pa.ptr, mode=mode, words_per_tile=T)
            lr, pr = _reference(hg, src, mode=mode, words_per_tile=T)
            assert np.array_equal(lv.numpy(), lr), f"graph {gi} call {call}: levels"
            assert np.array_equal(pa.numpy(), pr), f"graph {gi} call {call}: parents"
            assert st.ms_total > 0
            if mode == xb.XB_MODE_PLAIN:
                assert st.kernel_launches[10] > 0  # plain push levels ran
            else:
                assert st.tiles >= 1
            lv.free()
            pa.free()


def test_noalloc_deep_tile_and_duplicates():
    """A tile deeper than the 8-bit level code takes the plain path inside the
    call; duplicate sources share a frontier entry."""
    ro, ci = tg.path(300)
    hg = xb.HGraph(ro, ci)
    src = np.array([0, 299, 150, 150, 0], dtype=np.int32)
    lv, pa = _tables(hg, src.size)
    st = xb.msbfs_run_noalloc(hg, src, lv.ptr, pa.ptr, mode=xb.XB_MODE_BATCHED)
    lr, pr = _reference(hg, src, mode=xb.XB_MODE_PLAIN)
    assert np.array_equal(lv.numpy(), lr) and np.array_equal(pa.numpy(), pr)
    assert st.fallback_sources == src.size


def test_noalloc_limits():
    # Batched tiles of a 1M-vertex graph do not fit the 64 MiB array; the plain
    # engine's queues (16 B per vertex) do, and XB_MODE_AUTO takes them.
    g = xb.gen_rmat(20, 2, 0.57, 0.19, 0.19, seed=34)
    hg = xb.HGraph(g.row_offsets, g.col_indices)
    src = xb.sample_sources(g, 3, seed=35)
    lv, pa = _tables(hg, src.size)
    with pytest.raises(xb.XbError) as ei:
        xb.msbfs_run_noalloc(hg, src, lv.ptr, pa.ptr, mode=xb.XB_MODE_BATCHED)
    assert ei.value.code == xb.XB_ERR_BUDGET
    st = xb.msbfs_run_noalloc(hg, src, lv.ptr, pa.ptr, mode=xb.XB_MODE_AUTO)
    assert st.kernel_launches[10] > 0 and st.tiles == 0
    lr, pr = _reference(hg, src, mode=xb.XB_MODE_PLAIN)
    assert np.array_equal(lv.numpy(), lr) and np.array_equal(pa.numpy(), pr)
    # Other cut-offs than the graph's layout: refused until a run rebuilds it.
    with pytest.raises(xb.XbError) as ei:
        xb.msbfs_run_noalloc(hg, src, lv.ptr, pa.ptr, mode=xb.XB_MODE_AUTO, pull_cut_small=4)
    assert ei.value.code == xb.XB_ERR_UNSUPPORTED
    with pytest.raises(xb.XbError) as ei:
        xb.msbfs_run_noalloc(hg, src, 0, pa.ptr)
    assert ei.value.code == xb.XB_ERR_INVALID_ARG
    lib = xb.load_engine()
    for flag in (xb.XB_NO_LEVELS, xb.XB_PROFILE):
        o = xb.default_opts(flags=xb.XB_OUT_DEVICE | flag)
        rc = lib.xb_msbfs_run_noalloc(hg.handle, xb._ptr(src), src.size, xb._ptr(lv.ptr), xb._ptr(pa.ptr),
                                      ctypes.byref(o), None)
        assert rc == xb.XB_ERR_UNSUPPORTED
    with pytest.raises(xb.XbError) as ei:
        xb.msbfs_run_noalloc(hg, np.array([g.n], dtype=np.int32), lv.ptr, pa.ptr)
    assert ei.value.code == xb.XB_ERR_SOURCE_RANGE
    lv.free()
    pa.free()
    # A graph whose plain queues alone exceed the array: 5M vertices, no edges.
    n = 5_000_000
    hz = xb.HGraph(np.zeros(n + 1, dtype=np.int32), np.zeros(0, dtype=np.int32))
    lz, pz = _tables(hz, 1)
    with pytest.raises(xb.XbError) as ei:
        xb.msbfs_run_noalloc(hz, np.array([7], dtype=np.int32), lz.ptr, pz.ptr, mode=xb.XB_MODE_PLAIN)
    assert ei.value.code == xb.XB_ERR_BUDGET
