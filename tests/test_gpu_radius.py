"""Phase N parity: GPU neighbour lists equal the oracle's bit-for-bit."""

import numpy as np
import pytest

from oracle import c_oracle as C
from oracle import cnn_oracle as O
from gpu_util import gpu_csr

pytestmark = pytest.mark.gpu


def oracle_csr(X, r):
    X64 = np.asarray(X, dtype=np.float64)
    rp, col = C.neighbours_csr(X64, r, use_grid=X64.shape[1] <= 3)
    return rp, col.astype(np.int64)


def assert_same(X, r, precision):
    rp, col, st = gpu_csr(X, r, precision)
    rp0, col0 = oracle_csr(X, r)
    assert rp.shape == rp0.shape
    assert (rp == rp0).all(), "degrees differ"
    assert (col == col0).all(), "neighbour lists differ"
    # every unordered pair within r was computed; culling never adds work
    n = len(X)
    assert (len(col0) + n) // 2 <= st["pairs_evaluated"] <= st["tile_pairs_computed"] * 4096
    return st


@pytest.mark.parametrize("precision", ["f64", "f32_band"])
def test_blobs_2d_cfg(cuda, precision):
    from commonnnclustering_amd.synthetic import blobs_2d

    X, p = blobs_2d()
    assert_same(X, p["radius_cutoff"], precision)


@pytest.mark.parametrize("precision", ["f64", "f32_band"])
@pytest.mark.parametrize("n,d,dtype", [(1, 2, np.float64), (63, 3, np.float32), (64, 1, np.float64),
                                       (65, 2, np.float32), (500, 3, np.float64), (1000, 5, np.float32),
                                       (777, 8, np.float64), (600, 16, np.float32), (300, 64, np.float32),
                                       (257, 70, np.float64), (2000, 2, np.float32), (700, 17, np.float32),
                                       (400, 22, np.float32), (450, 24, np.float32), (500, 40, np.float32),
                                       (300, 78, np.float32), (300, 100, np.float32), (3000, 10, np.float32),
                                       (2500, 100, np.float64), (2100, 33, np.float32)])
def test_random_shapes(cuda, precision, n, d, dtype):
    rng = np.random.default_rng(n * 100 + d)
    X = np.concatenate([rng.normal(0, 1, size=(n // 2, d)), rng.uniform(-3, 3, size=(n - n // 2, d))])
    X = X.astype(dtype)
    # radius at a low percentile of pair distances
    D = np.sqrt(O.reduced_distances(X[:200], X[:200]))
    r = float(np.percentile(D[D > 0], 5)) if n > 1 else 1.0
    assert_same(X, r, precision)


@pytest.mark.parametrize("precision", ["f64", "f32_band"])
def test_boundary_pairs_exact(cuda, precision):
    """Pairs at exactly r and one ulp either side, in FP64 and FP32 inputs."""
    r = 1.5  # r*r = 2.25 exactly
    base = []
    for k, dist in enumerate([1.5, np.nextafter(1.5, 2.0), np.nextafter(1.5, 1.0)]):
        base += [[0.0, 10.0 * k], [dist, 10.0 * k]]
    # diagonal pair with d2 = 0.81 + 1.44 = 2.25 in exact arithmetic
    base += [[100.0, 100.0], [100.9, 101.2]]
    X = np.array(base, dtype=np.float64)
    st = assert_same(X, r, precision)
    X32 = X.astype(np.float32)
    assert_same(X32, r, precision)
    if precision == "f32_band":
        assert st["band_hits"] > 0


def test_band_hits_forced(cuda):
    """Many pairs placed inside the FP32 guard band must come out exact."""
    rng = np.random.default_rng(5)
    n = 4096
    d = 3
    A = rng.normal(size=(n // 2, d))
    u = rng.normal(size=(n // 2, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = 0.5
    eps = rng.uniform(-1e-6, 1e-6, size=(n // 2, 1))
    B = A + u * (r + eps)
    X = np.concatenate([A, B]).astype(np.float64)
    st = assert_same(X, r, "f32_band")
    assert st["band_hits"] > 100


@pytest.mark.parametrize("d", [1, 2, 3, 5, 8])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_band_offset_low_d(cuda, d, dtype):
    """Low-d pairs within ~1e-6 relative of r, far from the origin (f64 inputs
    lose bits to FP32 rounding, which widens the guard band): every band pair
    is settled in FP64 and the lists match the oracle.  (In float32 inputs at
    this offset the perturbation is below the coordinates' spacing, so few
    pairs fall in the band.)"""
    rng = np.random.default_rng(40 + d)
    m = 2048
    A = rng.normal(size=(m, d)) * 0.3 + 1000.0
    u = rng.normal(size=(m, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = 0.05
    B = A + u * (r * (1.0 + rng.uniform(-1e-6, 1e-6, size=(m, 1))))
    X = np.concatenate([A, B]).astype(dtype)
    st = assert_same(X, r, "f32_band")
    if dtype == np.float64:
        assert st["band_hits"] >= m // 4


@pytest.mark.parametrize("d", [2, 3])
def test_wide_tiles_low_d(cuda, d):
    """Tiles whose points span ~1e6 radii (a sparse far field around a dense
    core, float64 and float32 inputs): lists match the oracle."""
    rng = np.random.default_rng(70 + d)
    core = rng.normal(size=(300, d)) * 2e-3
    far = rng.uniform(-1e3, 1e3, size=(90, d))
    X = np.concatenate([core, far, far[:5] + 4e-4])
    rng.shuffle(X)
    assert_same(X, 1e-3, "f32_band")
    assert_same(X.astype(np.float32), 1e-3, "f32_band")


def test_duplicates_and_nonfinite(cuda):
    X = np.array([[0, 0], [0, 0], [np.nan, 0], [np.inf, 1], [0.1, 0.1], [-np.inf, 5], [0.2, -0.1]],
                 dtype=np.float64)
    for precision in ("f64", "f32_band"):
        rp, col, _ = gpu_csr(X, 0.5, precision)
        rp0, col0 = O.neighbourhoods_csr(X, 0.5)
        assert (rp == rp0).all() and (col == col0).all()


def test_empty(cuda):
    X = np.zeros((0, 3), dtype=np.float32)
    rp, col, st = gpu_csr(X, 1.0)
    assert rp.tolist() == [0] and col.size == 0


@pytest.mark.parametrize("d,dtype", [(16, np.float32), (64, np.float32), (33, np.float64)])
def test_gram_band_hits_forced(cuda, d, dtype):
    """High-d pairs placed within ~1e-6 relative of r: the MFMA Gram form must
    send them to the FP64 recheck and still match the oracle exactly."""
    rng = np.random.default_rng(d)
    m = 1500
    A = rng.normal(size=(m, d)) * 3.0 + 5.0   # offset: large norms relative to r
    u = rng.normal(size=(m, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = 2.0
    eps = rng.uniform(-1e-5, 1e-5, size=(m, 1))
    B = A + u * (r + eps)
    X = np.concatenate([A, B]).astype(dtype)
    st = assert_same(X, r, "f32_band")
    assert st["band_hits"] >= m // 2


def _clustered(n, d, seed):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, 4.0, size=(8, d))
    X = centres[rng.integers(0, 8, size=n)] + rng.normal(size=(n, d))
    D = np.sqrt(O.reduced_distances(X[:300], X[:300]))
    return X, float(np.percentile(D[D > 0], 3))


@pytest.mark.parametrize("d", [20, 64])
def test_gram_many_column_chunks(cuda, d):
    """More than one column chunk of the dense MFMA kernel (128 tiles each) and
    a ragged last tile."""
    X, r = _clustered(9000 + 37, d, d)
    assert_same(X.astype(np.float32), r, "f32_band")


@pytest.mark.parametrize("precision", ["f64", "f32_band"])
@pytest.mark.parametrize("d", [3, 12, 24])
def test_row_tile_batches(cuda, monkeypatch, precision, d):
    """A small mask budget splits the distance pass into row-tile batches (each
    batch evaluates all its tile pairs, no mirror writes); lists are unchanged."""
    X, r = _clustered(3000 + 11, d, 100 + d)
    monkeypatch.setenv("CNN_HIP_MASK_BUDGET", str(512 * 47 * 5))  # ~5 row tiles of 47 per batch
    st = assert_same(X.astype(np.float32), r, precision)
    assert st["distance_launches"] > 1


@pytest.mark.parametrize("precision", ["f64", "f32_band"])
def test_tile_pair_lists_searched_again(cuda, monkeypatch, precision):
    """Row tiles whose candidate lists overflow the first search pass's
    bucket (forced to 4 entries) are searched again in place; lists are
    unchanged."""
    X, r = _clustered(6000 + 5, 3, 7)
    monkeypatch.setenv("CNN_HIP_PAIR_BUCKET", "4")
    assert_same(X.astype(np.float32), r, precision)


@pytest.mark.parametrize("d", [16, 64])
def test_gram_band_redo_path(cuda, monkeypatch, d):
    """With room for a single band record, the other guard-band tile pairs are
    recomputed whole and settled by the redo kernel; lists stay exact."""
    monkeypatch.setenv("CNN_HIP_BAND_RECORDS", "1")
    rng = np.random.default_rng(d + 1)
    m = 800
    A = rng.normal(size=(m, d)) * 3.0 + 5.0
    u = rng.normal(size=(m, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = 2.0
    B = A + u * (r + rng.uniform(-1e-5, 1e-5, size=(m, 1)))
    st = assert_same(np.concatenate([A, B]).astype(np.float32), r, "f32_band")
    assert st["band_hits"] >= m // 2


def test_gram_huge_and_nonfinite(cuda):
    """MFMA form with points whose FP32 Gram terms overflow (|x| ~ 1e19, exact
    duplicates among them) and with inf / NaN coordinates: their tile pairs are
    settled in FP64, lists equal the oracle's."""
    rng = np.random.default_rng(11)
    d = 16
    X = rng.normal(size=(300, d))
    X[10] = 1e19
    X[11] = 1e19            # duplicate of 10: distance 0
    X[12] = -1e19
    X[70, 3] = np.inf
    X[71, 0] = np.nan
    X[72] = X[5] * 3e18
    for dtype in (np.float64, np.float32):
        Xc = X.astype(dtype)
        rp, col, _ = gpu_csr(Xc, 2.5, "f32_band")
        rp0, col0 = O.neighbourhoods_csr(Xc.astype(np.float64), 2.5)
        assert (rp == rp0).all() and (col == col0).all()


def _tiles_upper(n):
    T = (n + 63) // 64
    return T * (T + 1) // 2


@pytest.mark.parametrize("precision", ["f64", "f32_band"])
@pytest.mark.parametrize("d", [12, 24, 64, 100])
def test_dense_pivot_culling(cuda, precision, d):
    """Dense layout (d > 8, nearest-pivot order) on separated clusters: the
    pivot bounds cull most tile pairs and the lists stay exact."""
    X, r = _clustered(6000 + d, d, 300 + d)
    st = assert_same(X.astype(np.float32), r, precision)
    assert st["tile_pairs_computed"] < _tiles_upper(len(X)) // 2, st


@pytest.mark.parametrize("precision", ["f64", "f32_band"])
@pytest.mark.parametrize("d", [16, 40])
def test_dense_culling_tight_bound(cuda, precision, d):
    """Two groups of points on one axis whose closest pair lies at exactly r:
    every pivot lies on the axis, so the pivot bound of the tile pair holding
    that pair equals r, and the pair (a neighbour: s = r^2) must not be culled.
    Pairs one step past r across the gap are culled or rejected."""
    m = 1100
    step = 1.0 / 1024
    a = np.arange(m) * step                      # 0 .. (m-1)/1024
    r = 1.5                                      # r^2 = 2.25 exactly
    b = a[-1] + r + np.arange(m) * step          # first point exactly r past the last of a
    X = np.zeros((2 * m, d))
    X[:m, 0] = a
    X[m:, 0] = b
    st = assert_same(X.astype(np.float32), r, precision)
    assert st["tile_pairs_computed"] < _tiles_upper(2 * m), st


def test_dense_culling_nonfinite(cuda):
    """Dense layout with NaN, inf and huge rows (their tiles' bounds cull
    nothing) and exact duplicates: lists equal the oracle's."""
    X, r = _clustered(4000 + 5, 20, 77)
    X[100] = np.nan
    X[200, 3] = np.inf
    X[300] = 1e19
    X[301] = 1e19
    X[2500] = -1e19
    X[900] = X[901]
    for precision in ("f64", "f32_band"):
        for dtype in (np.float64, np.float32):
            assert_same(X.astype(dtype), r, precision)


@pytest.mark.parametrize("src_precision,precision", [("f32_band", "f32_band"), ("f64", "f32_band"),
                                                     ("f32_band", "f64"), ("f64", "f64")])
@pytest.mark.parametrize("d,n,sparse", [(24, 5000, False), (3, 6000, False), (2, 3000, True)])
def test_rethreshold_lists_exact(cuda, src_precision, precision, d, n, sparse):
    """A graph re-thresholded to smaller radii holds exactly the oracle's lists
    there: rebuilt from its working order (graphs that keep only their words,
    and d > 8) or filtered in FP64 (graphs that hold lists)."""
    import torch

    from commonnnclustering_amd import engine

    X, r = _clustered(n, d, 500 + d)
    if sparse:
        r *= 0.3  # few neighbours per word: the graph holds lists
    X = X.astype(np.float32)
    g = engine.build_graph(torch.from_numpy(X).cuda(), r * r, src_precision)
    try:
        for f in (0.9, 0.6):
            r2 = (f * r) ** 2
            g2 = g.rethreshold(r2, precision=precision)
            try:
                rp, col = g2.to_csr()
                rp0, col0 = oracle_csr(X, f * r)
                assert (rp.cpu().numpy() == rp0).all() and (col.cpu().numpy() == col0).all(), f
                assert g2.stats["nnz"] == len(col0)
            finally:
                g2.release()
    finally:
        g.release()


@pytest.mark.parametrize("d", [3, 20])
def test_graph_without_entries(cuda, d):
    """Only non-finite points: no point is a neighbour of any (not even of
    itself), the lists are empty and every label is noise."""
    from gpu_util import gpu_fit

    X = np.full((2500, d), np.nan)
    X[::7, 0] = np.inf
    for precision in ("f64", "f32_band"):
        rp, col, st = gpu_csr(X, 0.5, precision)
        assert (rp == 0).all() and col.size == 0 and st["nnz"] == 0
        lab, gs, ls = gpu_fit(X, 0.5, 2, precision=precision)
        assert (lab == 0).all() and ls["n_clusters"] == 0
