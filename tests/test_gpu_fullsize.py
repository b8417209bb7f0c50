"""Parity at the full sizes of the two headline workloads.

  md_basins_3d   n = 1,000,000, d = 3, float32, r = 0.25, c = 10, member_cutoff = 50
  features_64d   n = 200,000,  d = 64, float32, r = 1st pair-distance percentile, c = 20,
                 member_cutoff = 100

Both precision forms (plain FP64 and FP32 with the FP64 guard band) are fitted
through the C-ABI and compared with the C oracle (the restatement of
FitterExtBFS._fit, _fit.pyx:408-491, over NeighboursGetterExtBruteForce._get,
_types.pyx:1692-1710): every label, every neighbour row of the resident graph,
and 4096 seeded rows against the oracle's brute force with no prefilter.
These are the sizes at which the guard band, the row-tile batches, the tile
culling and the MFMA chunking all engage.
"""

import numpy as np
import pytest
import torch

from commonnnclustering_amd import _lib, engine
from commonnnclustering_amd.synthetic import features_64d, md_basins_3d
from oracle import c_oracle as C

pytestmark = pytest.mark.gpu


def _oracle(X, p, use_bf):
    r, c = p["radius_cutoff"], p["cnn_cutoff"]
    if use_bf:
        rp, col = C.neighbours_bf(X, r)
    else:
        rp, col = C.neighbours_csr(X, r)
    lab, _ = C.labels_from_csr(rp, col, c + 2, c + 1)
    C.sort_by_size(lab, p["member_cutoff"])
    return rp, col, lab


@pytest.fixture(scope="module")
def md3d():
    X, p = md_basins_3d()
    print(f"\n[fullsize] md_basins_3d oracle (n={len(X)}, {C.num_threads()} threads)", flush=True)
    rp, col, lab = _oracle(X, p, use_bf=False)
    return X, p, rp, col, lab


@pytest.fixture(scope="module")
def f64d():
    X, p = features_64d()
    print(f"\n[fullsize] features_64d oracle (n={len(X)}, {C.num_threads()} threads)", flush=True)
    rp, col, lab = _oracle(X, p, use_bf=True)
    return X, p, rp, col, lab


def _fit_and_compare(X, p, rp, col, want, precision):
    r, c, m = p["radius_cutoff"], p["cnn_cutoff"], p["member_cutoff"]
    Xd = torch.from_numpy(X).cuda()
    lab, g, gs, ls = engine.fit(Xd, r * r, c + 2, c + 1, precision=precision, member_cutoff=m, keep_graph=True)
    got = lab.cpu().numpy()
    mism = int((got != want).sum())
    print(f"[fullsize] n={len(X)} d={X.shape[1]} {precision}: label mismatches {mism}, "
          f"nnz {gs['nnz']} (oracle {int(rp[-1])}), band hits {gs['band_hits']}, flips {gs['band_flips']}, "
          f"max FP32 error {gs['max_abs_err_band']:.3g}", flush=True)
    assert mism == 0
    assert ls["n_clusters"] == len(np.unique(want[want > 0]))
    try:
        # the same graph labelled with the other similarity form: lists merged
        # (emitted on demand when the graph kept only its words)
        lab_m, st_m = g.labels(c + 2, c + 1, sim_mode=_lib.CNN_SIM_MERGE)
        assert st_m["sim_form"] == _lib.CNN_SIM_FORM_MERGE
        engine.sort_by_size(lab_m, m)
        assert np.array_equal(lab_m.cpu().numpy(), want), "merge form differs"
        del lab_m
        g_rp, g_col = g.to_csr()
        g_rp = g_rp.cpu().numpy()
        assert np.array_equal(g_rp, rp), "row offsets differ"
        g_col = g_col.cpu().numpy()
        assert g_col.dtype == col.dtype
        bad = np.flatnonzero(g_col != col)
        assert bad.size == 0, f"{bad.size} neighbour entries differ (first at {bad[:5]})"
        # seeded rows against the unfiltered brute force (no grid, no tiles)
        rows = np.random.default_rng(17).choice(len(X), size=4096, replace=False)
        brp, bcol = C.neighbours_bf(X, r, rows=rows)
        for t, i in enumerate(rows):
            assert np.array_equal(g_col[g_rp[i]:g_rp[i + 1]], bcol[brp[t]:brp[t + 1]]), f"row {i}"
    finally:
        g.release()
    return gs, ls


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision", ["f64", "f32_band"])
def test_md_basins_3d_full_size(cuda, md3d, precision):
    X, p, rp, col, want = md3d
    gs, ls = _fit_and_compare(X, p, rp, col, want, precision)
    assert gs["layout"] == 1 and gs["distance_launches"] == 1
    # one batch keeps the tile-pair words: similarities on words, no list emission in the fit
    assert ls["sim_form"] == _lib.CNN_SIM_FORM_TILE_WORDS
    assert gs["ms_emit"] == 0.0
    if precision == "f32_band":
        assert gs["band_hits"] > 0  # the guard band is exercised at this size


@pytest.mark.timeout(900)
def test_md_basins_3d_row_tile_batches(cuda, md3d, monkeypatch):
    """A mask budget that splits the build into row-tile batches: lists are
    emitted batch by batch, the words dropped, and similarities tested on
    per-row word lists built from the lists -- the same labels."""
    X, p, rp, col, want = md3d
    monkeypatch.setenv("CNN_HIP_MASK_BUDGET", str(200 << 20))
    r, c, m = p["radius_cutoff"], p["cnn_cutoff"], p["member_cutoff"]
    lab, _, gs, ls = engine.fit(torch.from_numpy(X).cuda(), r * r, c + 2, c + 1, precision="f32_band",
                                member_cutoff=m)
    print(f"[fullsize] md_basins_3d batches={gs['distance_launches']} sim_form={ls['sim_form']} "
          f"mismatches={int((lab.cpu().numpy() != want).sum())}", flush=True)
    assert gs["distance_launches"] > 1
    assert ls["sim_form"] in (_lib.CNN_SIM_FORM_ROW_WORDS, _lib.CNN_SIM_FORM_MERGE)
    assert np.array_equal(lab.cpu().numpy(), want)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision", ["f64", "f32_band"])
def test_features_64d_full_size(cuda, f64d, precision):
    X, p, rp, col, want = f64d
    gs, ls = _fit_and_compare(X, p, rp, col, want, precision)
    if precision == "f32_band":
        assert gs["band_hits"] > 0
