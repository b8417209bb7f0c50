"""Similarity forms: the sorted-list merge and the bitset (AND + popcount of
row words) give the oracle's labels in every mode, on resident graphs and on
caller lists, and a graph that keeps its words emits its lists only on demand.

Reference semantics: SimilarityCheckerExtSwitchContains._check /
SimilarityCheckerExtScreensorted._check (_types.pyx:2580-2610, 2654-2706)
inside FitterExtBFS._fit (_fit.pyx:408-491)."""

import numpy as np
import pytest
import torch

from commonnnclustering_amd import _lib, engine
from gpu_util import to_dev
from oracle import c_oracle as C

pytestmark = pytest.mark.gpu

AUTO, MERGE, BITSET = _lib.CNN_SIM_AUTO, _lib.CNN_SIM_MERGE, _lib.CNN_SIM_BITSET
F_MERGE, F_TILE, F_ROW = _lib.CNN_SIM_FORM_MERGE, _lib.CNN_SIM_FORM_TILE_WORDS, _lib.CNN_SIM_FORM_ROW_WORDS


def _data(seed, n, d, k=5, spread=4.0, scale=0.6, dtype=np.float32):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-spread, spread, size=(k, d))
    X = centres[rng.integers(0, k, size=n)] + rng.normal(size=(n, d)) * scale
    return X.astype(dtype)


CASES = [
    # seed, n, d, r, c
    (1, 4000, 2, 0.35, 5),
    (2, 6100, 3, 0.5, 12),
    (3, 3000, 2, 0.12, 1),     # low degree: lists kept, words dropped
    (4, 2500, 8, 1.6, 9),
    (5, 3300, 16, 3.4, 20),    # dense layout (pivot order)
    (6, 2000, 40, 5.0, 0),
    (7, 5000, 1, 0.02, 3),
    (8, 4500, 3, 0.9, 60),
]


@pytest.mark.parametrize("seed,n,d,r,c", CASES)
@pytest.mark.parametrize("precision", ["f64", "f32_band"])
def test_graph_labels_every_mode(cuda, seed, n, d, r, c, precision):
    X = _data(seed, n, d)
    rng = np.random.default_rng(seed)
    consider = (rng.uniform(size=n) > 0.1).astype(np.uint8)
    init = np.where(consider == 0, rng.integers(1, 4, size=n), 0).astype(np.int64)
    start = int(rng.integers(1, 9))
    rp, col = C.neighbours_csr(X, r, use_grid=d <= 3)
    want, _ = C.labels_from_csr(rp, col, c + 2, c + 1, consider=consider, current_start=start, labels=init.copy())
    g = engine.build_graph(to_dev(X), float(r) * float(r), precision)
    try:
        # words kept (spatial graphs, dense graphs after culling): no lists emitted
        kept_words = g.stats["ms_emit"] == 0.0
        forms = {}
        for mode in (AUTO, BITSET, MERGE, AUTO):
            lab = to_dev(init.copy())
            con = to_dev(consider.copy())
            _, st = g.labels(c + 2, c + 1, consider=con, current_start=start, labels=lab, sim_mode=mode)
            assert (lab.cpu().numpy() == want).all(), (mode, st)
            assert not con.cpu().numpy().any()
            forms.setdefault(mode, st["sim_form"])
        assert forms[MERGE] == F_MERGE
        assert forms[BITSET] in (F_TILE, F_ROW)
        if kept_words:
            assert forms[AUTO] == F_TILE and forms[BITSET] == F_TILE
        # the lists (emitted on demand when only words were kept) are the oracle's
        g_rp, g_col = g.to_csr()
        assert (g_rp.cpu().numpy() == rp).all() and (g_col.cpu().numpy() == col).all()
    finally:
        g.release()


@pytest.mark.parametrize("seed,n,d,r,c", CASES[:5])
def test_caller_lists_every_mode(cuda, seed, n, d, r, c):
    X = _data(seed, n, d, dtype=np.float64)
    rp, col = C.neighbours_csr(X, r, use_grid=d <= 3)
    want, _ = C.labels_from_csr(rp, col, c + 2, c + 1)
    # the automatic choice estimates the fill on every stride-th row (at least
    # every 16th, about 2^21 entries in all)
    stride = max(16, -(-int(rp[-1]) // (1 << 21)))
    rows = range(0, n, stride)
    fill = sum(rp[i + 1] - rp[i] for i in rows) / max(1, sum(len(np.unique(col[rp[i]:rp[i + 1]] >> 6))
                                                         for i in rows))
    for mode in (AUTO, MERGE, BITSET):
        lab, st = engine.labels_from_csr(to_dev(rp), to_dev(col), c + 2, c + 1, sim_mode=mode)
        assert (lab.cpu().numpy() == want).all(), mode
        if mode == MERGE:
            assert st["sim_form"] == F_MERGE
        elif mode == BITSET:
            assert st["sim_form"] == F_ROW
        else:
            assert st["sim_form"] == (F_ROW if fill >= 8 else F_MERGE), fill


def test_row_words_long_rows(cuda):
    """Word lists built from lists whose words straddle 64-entry windows (rows of
    a few hundred entries over scattered tiles), checked through labels that
    need almost every common neighbour (c close to the degrees)."""
    rng = np.random.default_rng(9)
    n = 3000
    rows = []
    for i in range(n):
        base = rng.choice(n, size=int(rng.integers(100, 400)), replace=False)
        rows.append(np.unique(np.concatenate([base, [i]])))
    # symmetrise
    pairs = {(i, int(j)) for i, r in enumerate(rows) for j in r}
    pairs |= {(j, i) for i, j in pairs}
    lists = [[] for _ in range(n)]
    for i, j in pairs:
        lists[i].append(j)
    lists = [np.array(sorted(x), dtype=np.int32) for x in lists]
    rp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum([len(x) for x in lists], out=rp[1:])
    col = np.concatenate(lists)
    for c in (1, 20, 40, 80):
        want, _ = C.labels_from_csr(rp, col, c + 2, c + 1)
        for mode in (MERGE, BITSET):
            lab, st = engine.labels_from_csr(to_dev(rp), to_dev(col), c + 2, c + 1, sim_mode=mode)
            assert (lab.cpu().numpy() == want).all(), (c, mode)


def test_fit_keeps_words_and_skips_emission(cuda):
    """A dense spatial fit tests similarities on the tile-pair words and writes
    no lists; a kept graph emits them when exported or re-thresholded."""
    X = _data(21, 60000, 3, k=6, spread=3.0, scale=0.5)
    r, c, m = 0.3, 10, 20
    want = C.fit(X, r, c, member_cutoff=m)
    lab, g, gs, ls = engine.fit(to_dev(X), r * r, c + 2, c + 1, precision="f32_band", member_cutoff=m,
                                keep_graph=True)
    try:
        assert (lab.cpu().numpy() == want).all()
        # words are kept when they take no more memory than the lists would
        assert 512 * gs["tile_pairs"] <= 4 * gs["nnz"], (gs["tile_pairs"], gs["nnz"])
        assert ls["sim_form"] == F_TILE
        assert gs["ms_emit"] == 0.0
        rp, col = C.neighbours_csr(X, r)
        g_rp, g_col = g.to_csr()
        assert (g_rp.cpu().numpy() == rp).all() and (g_col.cpu().numpy() == col).all()
        assert g.stats["ms_emit"] > 0.0
        r2 = 0.2
        g2 = g.rethreshold(r2 * r2)
        try:
            lab2, _ = g2.fit(c + 2, c + 1, member_cutoff=m)
            assert (lab2.cpu().numpy() == C.fit(X, r2, c, member_cutoff=m)).all()
        finally:
            g2.release()
    finally:
        g.release()
