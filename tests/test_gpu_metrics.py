"""The reference's other metrics on the GPU: the sqrt Euclidean metric
(decided through the reduced-sum threshold) and precomputed distance matrices
give the oracle's neighbour lists and labels."""

import math

import numpy as np
import pytest
import torch

from oracle import cnn_oracle as O
from commonnnclustering_amd import Clustering, engine

pytestmark = pytest.mark.gpu


def _ring_data(seed, d=2):
    """Clusters plus points placed on spheres of radius r around other points,
    so many pairs sit at the threshold."""
    rng = np.random.default_rng(seed)
    r = 0.45
    base = np.concatenate([rng.normal(c, 0.4, size=(300, d)) for c in rng.uniform(-3, 3, size=(3, d))])
    u = rng.normal(size=(400, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    ring = base[rng.integers(0, len(base), size=400)] + r * u
    return np.concatenate([base, ring]), r


@pytest.mark.parametrize("precision", ["f64", "f32_band"])
@pytest.mark.parametrize("d", [2, 3, 16])
def test_sqrt_metric_lists(cuda, precision, d):
    X, r = _ring_data(d, d)
    if d == 16:
        r = 2.2
    if precision == "f32_band":
        X = X.astype(np.float32)
    T = engine.sqrt_decision_threshold(r)
    rp, col, _ = engine.radius_csr(torch.from_numpy(X).cuda(), T, precision)
    rp0, col0 = O.neighbourhoods_csr(X.astype(np.float64), r, metric="euclidean")
    assert (rp.cpu().numpy() == rp0).all()
    assert (col.cpu().numpy().astype(np.int64) == col0).all()


@pytest.mark.parametrize("recipe", ["coordinates", "coordinates_f32"])
def test_sqrt_metric_fit(cuda, recipe):
    X, r = _ring_data(7)
    cl = Clustering(X, recipe=recipe, fitter__ngetter__dgetter__metric="euclidean")
    cl.fit(r, 3, member_cutoff=5, v=False)
    want, _ = O.fit(X, r, 3, member_cutoff=5, metric="euclidean")
    assert (cl.labels == want).all()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_precomputed_fit(cuda, dtype):
    rng = np.random.default_rng(5)
    P = np.concatenate([rng.normal(c, 0.5, size=(400, 2)) for c in ([0, 0], [3, 1], [-2, 2])])
    D = np.sqrt(O.reduced_distances(P, P)).astype(dtype)
    r = 0.35
    cl = Clustering(D, recipe="distances")
    cl.fit(r, 4, member_cutoff=5, v=False)
    want, _ = O.fit(D, r, 4, member_cutoff=5, metric="precomputed")
    assert (cl.labels == want).all()
    rp, col, asym = engine.precomputed_csr(torch.from_numpy(np.ascontiguousarray(D)).cuda(), r)
    rp0, col0 = O.neighbourhoods_csr(D, r, metric="precomputed")
    assert asym == 0
    assert (rp.cpu().numpy() == rp0).all() and (col.cpu().numpy() == col0).all()


def test_precomputed_asymmetric_rejected(cuda):
    rng = np.random.default_rng(6)
    D = rng.uniform(0, 1, size=(130, 130))
    D = (D + D.T) / 2
    D[3, 100] = 0.0   # one directed pair within r that its mirror is not
    D[100, 3] = 0.99
    rp, col, asym = engine.precomputed_csr(torch.from_numpy(D).cuda(), 0.1)
    assert asym == 1  # the pair {3, 100}
    cl = Clustering(D, recipe="distances")
    with pytest.raises(NotImplementedError):
        cl.fit(0.1, 1, v=False)


def _periodic_metric(bounds):
    from commonnnclustering_amd._types import MetricExtEuclideanPeriodicReduced
    return (MetricExtEuclideanPeriodicReduced, (np.asarray(bounds, dtype=float),), {})


def test_periodic_dihedrals_fit(cuda):
    """The reference's periodic-metric case (test_procedures.py:175-215):
    backbone dihedrals in degrees, bounds (360, 360), fit(10, 15, member_cutoff=50)."""
    import os
    data = np.load(os.path.join(os.path.dirname(__file__), "golden", "backbone_dihedrals.npy"))
    cl = Clustering(data, fitter__getter__dgetter__metric=_periodic_metric([360, 360]))
    cl.fit(10, 15, member_cutoff=50, v=False)
    want, _ = O.fit(data, 10, 15, member_cutoff=50, metric="periodic", bounds=[360.0, 360.0])
    assert (cl.labels == want).all()
    assert len(set(want.tolist()) - {0}) >= 2


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_periodic_lists_beyond_one_period(cuda, dtype):
    rng = np.random.default_rng(9)
    X = np.concatenate([rng.normal(170, 12, size=(400, 3)), rng.uniform(-800, 800, size=(300, 3))])
    X[:, 2] = rng.normal(0, 3, size=len(X))      # a non-periodic axis
    X = X.astype(dtype)
    bounds = [360.0, 360.0, 0.0]
    g = engine.build_graph(torch.from_numpy(X).cuda(), 18.0 ** 2, "f32_band", bounds=bounds)
    try:
        rp, col = g.to_csr()
        with pytest.raises(Exception):
            g.rethreshold(10.0 ** 2)   # periodic graphs are rebuilt, not re-thresholded
    finally:
        g.release()
    rp0, col0 = O.neighbourhoods_csr(X.astype(np.float64), 18.0, metric="periodic", bounds=bounds)
    assert (rp.cpu().numpy() == rp0).all()
    assert (col.cpu().numpy().astype(np.int64) == col0).all()
