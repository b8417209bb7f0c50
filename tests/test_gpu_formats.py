"""Output formats on GPU fits: Clustering.to_dtrajs (cluster.pyx:1166-1188)
and Summary.to_DataFrame (report.py:55-88), checked against the oracle's labels."""

import numpy as np
import pandas as pd
import pytest

from commonnnclustering_amd import Clustering
from commonnnclustering_amd.report import Record
from oracle import c_oracle as C

pytestmark = pytest.mark.gpu


def _parts(seed=11):
    """Three 'trajectories' of uneven length sampled from the same 2-D blobs."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-6, 6, size=(3, 2))
    parts = []
    for n in (1500, 700, 2300):
        lab = rng.integers(0, 3, size=n)
        parts.append(centres[lab] + rng.normal(scale=0.7, size=(n, 2)))
    return parts


@pytest.mark.parametrize("recipe", ["coordinates", "coordinates_f32"])
def test_to_dtrajs_on_gpu_fit(cuda, recipe):
    parts = _parts()
    cl = Clustering(parts, recipe=recipe)
    assert cl._bundle._input_data.meta["edges"] == [len(p) for p in parts]
    cl.fit(0.5, 8, member_cutoff=20, v=False)
    X = np.concatenate(parts)
    if recipe == "coordinates_f32":
        X = X.astype(np.float32)
    want = C.fit(X, 0.5, 8, member_cutoff=20)
    assert (cl.labels == want).all()
    d = cl.to_dtrajs()
    assert [len(x) for x in d] == [len(p) for p in parts]
    edges = np.cumsum([0] + [len(p) for p in parts])
    for k, x in enumerate(d):
        assert np.array_equal(x, want[edges[k]:edges[k + 1]])


def test_summary_dataframe_on_gpu_fits(cuda):
    X = np.concatenate(_parts(5))
    cl = Clustering(X)
    grid = [(0.4, 5, 10), (0.5, 8, 20), (0.8, 15, 2)]
    for r, c, m in grid:
        cl.fit(r, c, member_cutoff=m, v=False)
    df = cl.summary.to_DataFrame()
    assert list(df.columns) == Record.__slots__
    assert len(df) == len(grid)
    assert df["n_points"].dtype == pd.Int64Dtype() and df["radius_cutoff"].dtype == np.float64
    assert df["n_clusters"].dtype == pd.Int64Dtype() and df["ratio_noise"].dtype == np.float64
    for row, (r, c, m) in zip(df.itertuples(index=False), grid):
        want = C.fit(X, r, c, member_cutoff=m)
        k = len(np.unique(want[want > 0]))
        assert row.n_points == len(X)
        assert row.radius_cutoff == r and row.cnn_cutoff == c and row.member_cutoff == m
        assert row.n_clusters == k
        sizes = np.bincount(want)[1:] if k else np.zeros(1)
        assert row.ratio_largest == pytest.approx(sizes.max() / len(X))
        assert row.ratio_noise == pytest.approx(float((want == 0).sum()) / len(X))
