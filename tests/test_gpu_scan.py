"""Parameter scans (cnn_fit_scan): labels for every (r, c) pair equal a fresh
fit at that pair; the graph is built once at the largest radius."""

import numpy as np
import pytest
import torch

from oracle import c_oracle as C
from commonnnclustering_amd import engine, synthetic

pytestmark = pytest.mark.gpu


def _subset(n, seed=11):
    X, p = synthetic.scan_2d()
    idx = np.sort(np.random.default_rng(seed).choice(X.shape[0], size=n, replace=False))
    return np.ascontiguousarray(X[idx]), p


@pytest.mark.parametrize("precision", ["f32_band", "f64"])
def test_scan_grid_matches_fresh_fits(cuda, precision):
    X, p = _subset(6000)
    radii, cs = list(p["radii"]), list(p["cnn_cutoffs"])
    lab, gs, ls = engine.fit_scan(torch.from_numpy(X).cuda(), radii, cs, precision=precision, member_cutoff=2)
    lab = lab.cpu().numpy()
    assert lab.shape == (len(radii), len(cs), X.shape[0])
    for ri, r in enumerate(radii):
        for ci, c in enumerate(cs):
            want = C.fit(X, r, c, member_cutoff=2)
            assert (lab[ri, ci] == want).all(), (r, c)
            assert ls[ri * len(cs) + ci]["n_clusters"] == int(want.max(initial=0))


def test_scan_unsorted_radii_and_duplicates(cuda):
    X, p = _subset(3000, seed=12)
    radii = [0.3, 0.8, 0.3, 0.15]
    cs = [5, 0, 20]
    lab, _, _ = engine.fit_scan(torch.from_numpy(X).cuda(), radii, cs, member_cutoff=5, max_clusters=3)
    lab = lab.cpu().numpy()
    for ri, r in enumerate(radii):
        for ci, c in enumerate(cs):
            want = C.fit(X, r, c, member_cutoff=5, max_clusters=3)
            assert (lab[ri, ci] == want).all(), (r, c)


def test_scan_equals_fit_on_gpu(cuda):
    X, p = _subset(4000, seed=13)
    Xd = torch.from_numpy(X).cuda()
    lab, _, _ = engine.fit_scan(Xd, [0.5, 0.25], [3, 10], precision="f32_band")
    for ri, r in enumerate([0.5, 0.25]):
        for ci, c in enumerate([3, 10]):
            one, _, _, _ = engine.fit(Xd, r * r, c + 2, c + 1, precision="f32_band")
            assert torch.equal(lab[ri, ci], one)
