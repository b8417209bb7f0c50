"""Seeded synthetic workloads for tests and benchmarks (numpy default_rng).

  blobs_2d          n=2,000, d=2, float64: 4 Gaussian blobs (centres U[-10,10]^2,
                    std 1, 450 points each) + 200 uniform points in [-12,12]^2;
                    r=1.0, c=5, member_cutoff=10
  md_basins_3d      n=1,000,000, d=3, float32: 8 anisotropic basins (centres
                    U[-5,5]^3, per-axis std U[0.2,0.8], equal split) + 5% uniform
                    noise in [-6,6]^3; r=0.25, c=10, member_cutoff=50
  features_64d      n=200,000, d=64, float32: 20 clusters (centres 2*N(0,I),
                    std 1, sizes multinomial(n, Dirichlet(1))); r = 1st
                    percentile of FP64 distances over 1e5 random pairs;
                    c=20, member_cutoff=100
  scan_2d           n=50,000, d=2, float32: 6 blobs (std 0.5, centres
                    U[-8,8]^2) + 10% uniform noise in [-8,8]^2; 64 (r, c) pairs
  many_small        512 datasets, n log-uniform in [500, 20000], d in {2,3,8},
                    3-10 blobs (centres U[-10,10]^d, std 1); r at the 2nd
                    percentile of 1e5 pair distances, c=5
  serving_request   one request of the serving workload (n log-uniform in
                    [1e3, 3e5], d in {2,3,16,64}, 2-30 blobs + 5% noise)
  sklearn_tutorial  the six n=2000 datasets of the reference's scikit-learn
                    tutorial (scikit_learn_datasets.ipynb:205-322)
"""

from __future__ import annotations

import numpy as np


def blobs_2d(seed: int = 1, n_blob: int = 450, n_noise: int = 200):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-10, 10, size=(4, 2))
    parts = [rng.normal(c, 1.0, size=(n_blob, 2)) for c in centres]
    parts.append(rng.uniform(-12, 12, size=(n_noise, 2)))
    X = np.concatenate(parts).astype(np.float64)
    return X, {"radius_cutoff": 1.0, "cnn_cutoff": 5, "member_cutoff": 10}


def md_basins_3d(n: int = 1_000_000, seed: int = 2, noise_frac: float = 0.05):
    rng = np.random.default_rng(seed)
    n_noise = int(round(n * noise_frac))
    n_basin = n - n_noise
    sizes = [n_basin // 8 + (1 if b < n_basin % 8 else 0) for b in range(8)]
    centres = rng.uniform(-5, 5, size=(8, 3))
    stds = rng.uniform(0.2, 0.8, size=(8, 3))
    parts = [rng.normal(centres[b], stds[b], size=(sizes[b], 3)) for b in range(8)]
    parts.append(rng.uniform(-6, 6, size=(n_noise, 3)))
    X = np.concatenate(parts).astype(np.float32)
    return X, {"radius_cutoff": 0.25, "cnn_cutoff": 10, "member_cutoff": 50}


def pair_distance_percentile(X: np.ndarray, q: float, n_pairs: int = 100_000, rng=None) -> float:
    """q-th percentile (numpy linear) of FP64 Euclidean distances over random pairs i != j."""
    if rng is None:
        rng = np.random.default_rng(0)
    n = X.shape[0]
    i = rng.integers(0, n, size=n_pairs)
    j = rng.integers(0, n - 1, size=n_pairs)
    j = j + (j >= i)
    diff = X[i].astype(np.float64) - X[j].astype(np.float64)
    dist = np.sqrt(np.sum(diff * diff, axis=1))
    return float(np.percentile(dist, q))


def features_64d(n: int = 200_000, d: int = 64, seed: int = 3, k: int = 20):
    rng = np.random.default_rng(seed)
    weights = rng.dirichlet(np.ones(k))
    sizes = rng.multinomial(n, weights)
    centres = 2.0 * rng.normal(size=(k, d))
    parts = [rng.normal(centres[c], 1.0, size=(sizes[c], d)) for c in range(k)]
    X = np.concatenate(parts).astype(np.float32)
    r = pair_distance_percentile(X, 1.0, rng=rng)
    return X, {"radius_cutoff": r, "cnn_cutoff": 20, "member_cutoff": 100}


def scan_2d(n: int = 50_000, seed: int = 4, noise_frac: float = 0.10):
    rng = np.random.default_rng(seed)
    n_noise = int(round(n * noise_frac))
    n_blob = n - n_noise
    sizes = [n_blob // 6 + (1 if b < n_blob % 6 else 0) for b in range(6)]
    centres = rng.uniform(-8, 8, size=(6, 2))
    parts = [rng.normal(centres[b], 0.5, size=(sizes[b], 2)) for b in range(6)]
    parts.append(rng.uniform(-8, 8, size=(n_noise, 2)))
    X = np.concatenate(parts).astype(np.float32)
    radii = np.geomspace(0.1, 0.8, 8)
    cs = [1, 2, 5, 10, 20, 40, 80, 160]
    return X, {"radii": radii, "cnn_cutoffs": cs}


def one_small_dataset(seed: int, n_range=(500, 20_000), dims=(2, 3, 8)):
    rng = np.random.default_rng(seed)
    n = int(round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1])))))
    d = int(rng.choice(dims))
    k = int(rng.integers(3, 11))
    centres = rng.uniform(-10, 10, size=(k, d))
    lab = rng.integers(0, k, size=n)
    X = (centres[lab] + rng.normal(size=(n, d))).astype(np.float32)
    r = pair_distance_percentile(X, 2.0, rng=rng)
    return X, {"radius_cutoff": r, "cnn_cutoff": 5, "member_cutoff": 2}


def many_small(count: int = 512, base_seed: int = 4000):
    return [one_small_dataset(base_seed + i) for i in range(count)]


def serving_request(seed: int, n_range=(1_000, 300_000), dims=(2, 3, 16, 64)):
    rng = np.random.default_rng(seed)
    n = int(round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1])))))
    d = int(rng.choice(dims))
    k = int(rng.integers(2, 31))
    n_noise = int(round(0.05 * n))
    spread = 10.0 if d <= 3 else 4.0
    centres = rng.uniform(-spread, spread, size=(k, d))
    lab = rng.integers(0, k, size=n - n_noise)
    body = centres[lab] + rng.normal(size=(n - n_noise, d))
    noise = rng.uniform(-spread - 2, spread + 2, size=(n_noise, d))
    X = np.concatenate([body, noise]).astype(np.float32)
    q = float(rng.uniform(0.5, 3.0))
    r = pair_distance_percentile(X, q, rng=rng)
    c = int(rng.integers(2, 51))
    return X, {"radius_cutoff": r, "cnn_cutoff": c, "member_cutoff": 2}


def sklearn_tutorial():
    """The six n=2000 datasets of scikit_learn_datasets.ipynb (cells 6-8, :205-322),
    regenerated with the same global seed and generator calls, by name."""
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
    return {"circles": circles, "moons": moons, "varied": varied, "aniso": aniso, "blobs": blobs, "None": none}
