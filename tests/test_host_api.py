"""Host-side API pieces that need no GPU: trajectory splitting of labels and
the typed summary table."""

import numpy as np
import pandas as pd

from commonnnclustering_amd import Clustering
from commonnnclustering_amd.report import Record, Summary


def test_to_dtrajs_follows_input_parts():
    parts = [np.zeros((3, 2)), np.ones((4, 2)), np.full((2, 2), 2.0)]
    cl = Clustering(parts)
    assert cl._bundle._input_data.meta["edges"] == [3, 4, 2]
    cl.labels = np.arange(9)
    d = cl.to_dtrajs()
    assert [x.tolist() for x in d] == [[0, 1, 2], [3, 4, 5, 6], [7, 8]]


def test_to_dtrajs_without_labels_or_edges():
    cl = Clustering(np.zeros((5, 2)))
    assert cl.to_dtrajs() == []
    cl._bundle._input_data.meta.pop("edges", None)
    cl.labels = [1, 1, 0, 2, 2]
    d = cl.to_dtrajs()
    assert len(d) == 1 and d[0].tolist() == [1, 1, 0, 2, 2]


def test_summary_dataframe_types():
    s = Summary()
    s.append(Record(100, 0.5, 2, 10, None, 3, 0.4, 0.1, 0.02))
    s.append(Record(200, 0.25, 5, None, 4, 1, 0.9, 0.05, 0.01))
    df = s.to_DataFrame()
    assert list(df.columns) == Record.__slots__
    assert df["n_points"].dtype == pd.Int64Dtype() and df["radius_cutoff"].dtype == np.float64
    assert df["member_cutoff"].isna().tolist() == [False, True]
    assert df["n_clusters"].tolist() == [3, 1]


def test_label_counts_match_counter():
    """to_set and the Record statistics come from label_counts: the same
    values as the Counter / set over the labels that the reference uses
    (cluster.pyx:366-390, _types.pyx:200-210)."""
    from collections import Counter

    from commonnnclustering_amd._types import Labels, label_counts

    rng = np.random.default_rng(3)
    for lab in (rng.integers(0, 40, 5000), rng.integers(-3, 3, 100), np.array([7]), np.zeros(0, dtype=np.int64),
                rng.integers(0, 2 ** 40, 300)):
        uniq, counts = label_counts(lab)
        c = Counter(lab.tolist())
        assert dict(zip(uniq.tolist(), counts.tolist())) == dict(c)
        assert Labels(lab).to_set() == set(c)
