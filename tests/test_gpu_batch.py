"""Batched fits (cnn_fit_batch): every dataset's labels equal a fit of that
dataset alone, with its own radius and cutoffs, through both the packed path
(1 <= d <= 8: one working order for all datasets) and the per-dataset path."""

import numpy as np
import pytest
import torch

from oracle import c_oracle as C
from commonnnclustering_amd import engine, synthetic

pytestmark = pytest.mark.gpu

def _stack(sets):
    X = np.concatenate([x for x, _ in sets]).astype(np.float32)
    off = np.zeros(len(sets) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(x) for x, _ in sets])
    return X, off


def _oracle_each(sets, member_cutoff=None, max_clusters=None):
    out = []
    for x, p in sets:
        if len(x) == 0:
            out.append(np.zeros(0, dtype=np.int64))
            continue
        out.append(C.fit(x.astype(np.float32), p["radius_cutoff"], p["cnn_cutoff"], member_cutoff=member_cutoff,
                         max_clusters=max_clusters))
    return out


def _run_batch(sets, precision, member_cutoff=None, max_clusters=None):
    X, off = _stack(sets)
    r2 = np.array([float(p["radius_cutoff"]) ** 2 for _, p in sets])
    c = np.array([int(p["cnn_cutoff"]) for _, p in sets])
    Xd = torch.from_numpy(X).cuda()
    lab, gs, ls = engine.fit_batch(Xd, off, r2, c + 2, c + 1, precision=precision, member_cutoff=member_cutoff,
                                   max_clusters=max_clusters)
    lab = lab.cpu().numpy()
    return [lab[off[b]:off[b + 1]] for b in range(len(sets))], gs, ls


def _small_sets(d, count, seed):
    sets = []
    for i in range(count):
        x, p = synthetic.one_small_dataset(seed + 97 * i, n_range=(40, 3000), dims=(d,))
        p = dict(p)
        p["cnn_cutoff"] = int(np.random.default_rng(seed + i).integers(0, 8))
        sets.append((x, p))
    return sets


@pytest.mark.parametrize("precision", ["f64", "f32_band"])
@pytest.mark.parametrize("d", [2, 3, 8])
def test_packed_batch_matches_single_fits(cuda, d, precision):
    sets = _small_sets(d, 24, 100 * d)
    want = _oracle_each(sets, member_cutoff=2)
    got, gs, ls = _run_batch(sets, precision, member_cutoff=2)
    assert gs["layout"] == 1
    for b, (g, w) in enumerate(zip(got, want)):
        assert (g == w).all(), f"dataset {b} (n={len(w)})"
    assert ls["n_clusters"] == sum(int(w.max(initial=0)) for w in want)


def test_packed_batch_edge_sizes(cuda):
    # empty datasets, single points, exactly one tile, one tile + 1
    rng = np.random.default_rng(7)
    sizes = [0, 1, 2, 63, 64, 65, 0, 200, 1, 128]
    sets = []
    for n in sizes:
        x = rng.normal(size=(n, 3)).astype(np.float32)
        sets.append((x, {"radius_cutoff": 0.8, "cnn_cutoff": 1}))
    want = _oracle_each(sets, member_cutoff=1)
    got, _, _ = _run_batch(sets, "f32_band", member_cutoff=1)
    for b, (g, w) in enumerate(zip(got, want)):
        assert (g == w).all(), f"dataset {b} (n={len(w)})"


def test_packed_batch_max_clusters_and_identical_copies(cuda):
    x, p = synthetic.blobs_2d()
    sets = [(x, p), (x, p), (x, dict(p, radius_cutoff=0.7))]
    want = _oracle_each(sets, member_cutoff=10, max_clusters=2)
    got, _, ls = _run_batch(sets, "f64", member_cutoff=10, max_clusters=2)
    for g, w in zip(got, want):
        assert (g == w).all()
    assert (got[0] == got[1]).all()
    assert ls["n_clusters"] == sum(int(w.max(initial=0)) for w in want)


def test_batch_per_dataset_path_high_d(cuda):
    sets = []
    for i in range(3):
        x, p = synthetic.one_small_dataset(9000 + i, n_range=(300, 2000), dims=(16,))
        sets.append((x, p))
    want = _oracle_each(sets, member_cutoff=2)
    got, gs, _ = _run_batch(sets, "f32_band", member_cutoff=2)
    for g, w in zip(got, want):
        assert (g == w).all()


def test_batch_of_one_equals_fit(cuda):
    x, p = synthetic.blobs_2d()
    got, _, _ = _run_batch([(x, p)], "f32_band", member_cutoff=10)
    Xd = torch.from_numpy(x.astype(np.float32)).cuda()
    r = p["radius_cutoff"]
    lab, _, _, _ = engine.fit(Xd, r * r, p["cnn_cutoff"] + 2, p["cnn_cutoff"] + 1, precision="f32_band",
                              member_cutoff=10)
    assert (got[0] == lab.cpu().numpy()).all()


@pytest.mark.parametrize("d", [2, 3])
def test_batch_many_sets_concurrent(cuda, d):
    """More datasets than worker streams, uneven sizes, empty ones included:
    every dataset's labels equal a fit of it alone."""
    sets = _small_sets(d, 30, 500 + d)
    sets[4] = (np.zeros((0, d), dtype=np.float32), sets[4][1])
    sets[17] = (sets[17][0][:1], sets[17][1])
    want = _oracle_each(sets, member_cutoff=2)
    for _ in range(2):  # the worker contexts are reused by the second call
        got, gs, ls = _run_batch(sets, "f32_band", member_cutoff=2)
        for b, (g, w) in enumerate(zip(got, want)):
            assert (g == w).all(), f"dataset {b} (n={len(w)})"
        assert ls["n_clusters"] == sum(int(w.max(initial=0)) for w in want)
