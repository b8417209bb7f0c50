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


def test_scan_runs_captured_on_words(cuda):
    """A graph that keeps its row words runs the scan as one captured graph
    (every pair reports the tile-word similarity form)."""
    X, p = _subset(6000, seed=14)
    lab, gs, ls = engine.fit_scan(torch.from_numpy(X).cuda(), [0.8, 0.4], [2, 9], precision="f32_band")
    assert gs["layout"] == 1
    assert all(s["sim_form"] == 2 for s in ls)
    for ri, r in enumerate([0.8, 0.4]):
        for ci, c in enumerate([2, 9]):
            assert (lab[ri, ci].cpu().numpy() == C.fit(X, r, c, member_cutoff=2)).all()


def test_scan_graph_replays(cuda):
    """The captured scan replayed several times, with eager work and
    synchronisations between the replays: every replay equals fresh eager fits."""
    X, p = _subset(8000, seed=15)
    Xd = torch.from_numpy(X).cuda()
    radii, cs = [0.7, 0.35, 0.2], [1, 6, 30]
    g = engine.build_graph(Xd, 0.7 * 0.7, "f32_band")
    try:
        want = {}
        for ri, r in enumerate(radii):
            for ci, c in enumerate(cs):
                want[ri, ci] = engine.fit(Xd, r * r, c + 2, c + 1, precision="f32_band", member_cutoff=3)[0]
        sc = g.scan(radii, cs, member_cutoff=3)
        try:
            for rep in range(4):
                sc.labels.zero_()
                other = engine.fit(Xd, 0.09, 7, 6, precision="f64")[0]  # eager work between replays
                torch.cuda.synchronize()
                lab, stats = sc.launch()
                for (ri, ci), w in want.items():
                    assert torch.equal(lab[ri, ci], w), (rep, ri, ci)
                    assert stats[ri * len(cs) + ci]["n_clusters"] == int(w.max().item())
                del other
        finally:
            sc.release()
    finally:
        g.release()


def test_scan_graph_needs_words(cuda):
    X, p = _subset(3000, seed=16)
    g = engine.build_graph(torch.from_numpy(X).cuda(), 0.02 * 0.02, "f64")  # sparse: lists only
    try:
        from commonnnclustering_amd._lib import CNNCapacityError

        with pytest.raises(CNNCapacityError):
            g.scan([0.02], [3])
    finally:
        g.release()


def test_full_scan_workload(cuda):
    """The 64-pair scan of the n = 5e4 workload against independent fits."""
    X, p = synthetic.scan_2d()
    Xd = torch.from_numpy(X).cuda()
    radii, cs = list(p["radii"]), list(p["cnn_cutoffs"])
    lab, gs, ls = engine.fit_scan(Xd, radii, cs, precision="f32_band")
    assert all(s["sim_form"] == 2 for s in ls)
    for ri, r in enumerate(radii):
        for ci, c in enumerate(cs):
            one = engine.fit(Xd, float(r) ** 2, c + 2, c + 1, precision="f32_band")[0]
            assert torch.equal(lab[ri, ci], one), (r, c)
    # two pairs against the oracle as well
    for ri, ci in ((7, 2), (3, 5)):
        assert (lab[ri, ci].cpu().numpy() == C.fit(X, radii[ri], cs[ci], member_cutoff=2)).all()


@pytest.mark.parametrize("precision", ["f32_band", "f64"])
def test_scan_captured_on_dense_words(cuda, precision):
    """A dense-layout graph (d = 24, nearest-pivot order, culled tile pairs)
    keeps its words as well: its scan runs captured, each radius's words cut
    exactly from the largest, and every pair equals the oracle's fit."""
    rng = np.random.default_rng(17)
    d, n = 24, 5000
    centres = rng.normal(0.0, 4.0, size=(6, d))
    X = (centres[rng.integers(0, 6, size=n)] + rng.normal(size=(n, d))).astype(np.float32)
    radii, cs = [6.0, 5.2, 4.5], [3, 15, 60]
    lab, gs, ls = engine.fit_scan(torch.from_numpy(X).cuda(), radii, cs, precision=precision, member_cutoff=4)
    assert gs["layout"] == 0
    assert all(s["sim_form"] == 2 for s in ls)
    for ri, r in enumerate(radii):
        for ci, c in enumerate(cs):
            assert (lab[ri, ci].cpu().numpy() == C.fit(X, r, c, member_cutoff=4)).all(), (r, c)
