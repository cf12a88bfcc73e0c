"""Autotuner (xb_autotune): coordinate descent over words per tile, alpha and
the pull cut-offs on a sample; the tuned options keep results bit-exact."""
import numpy as np
import pytest

import oracle
import xbmsbfs as xb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graph():
    g = xb.gen_rmat(14, 16, 0.57, 0.19, 0.19, seed=21)
    return g, xb.HGraph(g.row_offsets, g.col_indices)


def test_autotune_report_and_parity(graph):
    g, hg = graph
    sample = xb.sample_sources(g, 300, seed=22)
    eng = xb.XBlossomGpuEngine10(hg)
    rep = eng.Autotune(sample)
    d = rep.as_dict()
    assert len(d["trials"]) >= 20 and 1 <= d["passes"] <= 2
    assert 0 < d["best_ms_per_source"] <= d["default_ms_per_source"]
    t = eng.tunables
    assert t["words_per_tile"] in (1, 2, 4, 8, 16)
    assert t["pull_cut_small"] <= t["pull_cut_medium"] <= t["pull_cut_hub"]
    # The tuned options give the same tables.
    src = xb.sample_sources(g, 150, seed=23)
    lv, pa = eng.MultiSourceBFS(src, mode=xb.XB_MODE_BATCHED)
    lr, pr, _ = oracle.msbfs(g.row_offsets, g.col_indices, src, threads=8)
    assert np.array_equal(lv, lr) and np.array_equal(pa, pr)


def test_autotune_sample_wider_than_eight_words(graph):
    """A 700-source sample: the 16-word trial runs a 700-row tile (the
    sample buffer holds one tile of the widest trial)."""
    g, hg = graph
    sample = xb.sample_sources(g, 700, seed=24)
    best, rep = xb.autotune(hg, sample)
    widths = {t["words_per_tile"] for t in rep.as_dict()["trials"]}
    assert {8, 16} <= widths
    lv, pa = xb.XBlossomGpuEngine10(hg).MultiSourceBFS(sample, mode=xb.XB_MODE_BATCHED, **best)
    lr, pr, _ = oracle.msbfs(g.row_offsets, g.col_indices, sample, threads=8)
    assert np.array_equal(lv, lr) and np.array_equal(pa, pr)


def test_autotune_rejects_bad_samples(graph):
    g, hg = graph
    with pytest.raises(xb.XbError) as ei:
        xb.autotune(hg, np.array([], dtype=np.int32))
    assert ei.value.code == xb.XB_ERR_INVALID_ARG
    with pytest.raises(xb.XbError) as ei:
        xb.autotune(hg, np.array([0, g.n], dtype=np.int32))
    assert ei.value.code == xb.XB_ERR_SOURCE_RANGE
