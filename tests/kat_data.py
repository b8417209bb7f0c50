"""Known-answer inputs shared by the CPU and GPU tests."""

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kats.json")


def kats():
    with open(GOLDEN) as fh:
        return json.load(fh)


def sklearn_tutorial_datasets():
    """The six n=2000 datasets of scikit_learn_datasets.ipynb (cells 6-8, :205-322),
    regenerated with the same global seed and generator calls, in record order."""
    from sklearn import datasets
    from sklearn.preprocessing import StandardScaler

    state = np.random.get_state()
    try:
        np.random.seed(0)
        n = 2000
        circles, _ = datasets.make_circles(n_samples=n, factor=.5, noise=.05)
        circles = StandardScaler().fit_transform(circles)
        moons, _ = datasets.make_moons(n_samples=n, noise=.05)
        moons = StandardScaler().fit_transform(moons)
        blobs, _ = datasets.make_blobs(n_samples=n, random_state=8)
        blobs = StandardScaler().fit_transform(blobs)
        none = StandardScaler().fit_transform(np.random.rand(n, 2))
        X, _ = datasets.make_blobs(n_samples=n, random_state=170)
        aniso = StandardScaler().fit_transform(np.dot(X, [[0.6, -0.6], [-0.4, 0.8]]))
        varied, _ = datasets.make_blobs(n_samples=n, cluster_std=[1.0, 2.5, 0.5], random_state=170)
        varied = StandardScaler().fit_transform(varied)
    finally:
        np.random.set_state(state)
    by_name = {"circles": circles, "moons": moons, "varied": varied, "aniso": aniso, "blobs": blobs,
               "None": none}
    return by_name


def equalise_labels(labels):
    """Relabel clusters by order of first appearance, 0 kept (test_procedures.py:20-42)."""
    out = np.array(labels, copy=True)
    mapping = {}
    for i, l in enumerate(out):
        if l == 0:
            continue
        if l not in mapping:
            mapping[l] = len(mapping) + 1
        out[i] = mapping[l]
    return out
