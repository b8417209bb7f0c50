"""Pins the CPU oracle (oracle/) against the reference's known answers and
against mathematics independent of the oracle's own formulas."""

from collections import Counter
from fractions import Fraction

import numpy as np
import pytest

from oracle import cnn_oracle as O
from oracle import c_oracle as C
from kat_data import equalise_labels, kats, sklearn_tutorial_datasets

K = kats()


# ------------------------------------------------------------------ metric

@pytest.mark.parametrize("d", [1, 2, 3, 7, 16, 64])
def test_metric_matches_scalar_loop(d):
    rng = np.random.default_rng(d)
    A = rng.normal(size=(40, d)) * rng.uniform(0.1, 100)
    B = rng.normal(size=(30, d))
    D = O.reduced_distances(A, B)
    for i in range(0, 40, 7):
        for j in range(0, 30, 5):
            # literal IEEE-double loop: total = total + (a-b)*(a-b)
            t = 0.0
            for k in range(d):
                diff = float(A[i, k]) - float(B[j, k])
                t = t + diff * diff
            assert D[i, j] == t  # bit-identical


def test_metric_error_bound_against_exact():
    rng = np.random.default_rng(7)
    u = 2.0 ** -53
    for d in (2, 3, 64):
        A = rng.normal(size=(20, d))
        D = O.reduced_distances(A, A)
        for i in range(20):
            for j in range(20):
                exact = O.exact_reduced_distance(A[i], A[j])
                err = abs(Fraction(D[i, j]) - exact)
                assert err <= Fraction(O.gamma(2 * d, u)) * exact


def test_metric_kat_points_vs_closed_form():
    m = K["metric"]
    X = np.array(m["data"], dtype=np.float64)
    Y = np.array(m["other"], dtype=np.float64)
    D = O.reduced_distances(Y, X)
    for i in range(len(Y)):
        for j in range(len(X)):
            ref = sum((a - b) ** 2 for a, b in zip(Y[i], X[j]))
            np.testing.assert_approx_equal(D[i, j], ref, significant=m["significant"])


def test_adjust_radius_is_square():
    for r in (0.1, 0.25, 1.5, 9.99, 1e-300, 1e200):
        assert O.adjust_radius(r) == r * r


# ------------------------------------------------------------------ neighbourhoods

def test_boundary_inclusive_and_one_ulp_beyond():
    r = 1.5  # r*r = 2.25 exactly
    far = np.nextafter(1.5, 2.0)
    X = np.array([[0.0, 0.0], [1.5, 0.0], [0.0, far], [-1.5, 0.0]])
    N = O.neighbourhoods(X, r)
    assert N[0].tolist() == [0, 1, 3]       # exactly at r: neighbour (<=)
    assert 2 not in N[0].tolist()           # one ulp beyond: not
    assert N[2].tolist() == [2]


def test_neighbourhoods_symmetric_and_self():
    rng = np.random.default_rng(3)
    X = rng.normal(size=(300, 3))
    N = O.neighbourhoods(X, 0.6)
    for i, row in enumerate(N):
        assert i in row.tolist()
        assert np.all(np.diff(row) > 0)
        for j in row:
            assert i in N[j].tolist()


# ------------------------------------------------------------------ similarity

@pytest.mark.parametrize("case", K["similarity"]["cases"])
def test_similarity_kat(case):
    assert O.check_similarity(case["a"], case["b"], case["c"]) == case["expected"]
    if case["sorted"]:
        a = np.array(case["a"], dtype=np.int32)
        b = np.array(case["b"], dtype=np.int32)
        assert bool(C.lib().oracle_check_similarity(a, len(a), b, len(b), case["c"])) == case["expected"]


# ------------------------------------------------------------------ sort_by_size

@pytest.mark.parametrize("case", K["sort_by_size"]["cases"])
def test_sort_by_size_kat(case):
    for fn in (O.sort_by_size, O.sort_by_size_np):
        lab = np.array(K["sort_by_size"]["labels"], dtype=np.int64)
        fn(lab, **case["kwargs"])
        assert lab.tolist() == case["expected"]
    lab = np.array(K["sort_by_size"]["labels"], dtype=np.int64)
    kw = case["kwargs"]
    C.sort_by_size(lab, kw.get("member_cutoff"), kw.get("max_clusters"))
    assert lab.tolist() == case["expected"]


def test_sort_by_size_ties_follow_counter_most_common():
    rng = np.random.default_rng(11)
    for trial in range(30):
        lab = rng.integers(0, 12, size=rng.integers(1, 80)).astype(np.int64) * rng.integers(1, 1000)
        # Counter.most_common is the definition of the order (stable on first occurrence)
        freq = Counter(lab.tolist())
        freq.pop(0, None)
        order = [k for k, _ in freq.most_common()]
        expect = lab.copy()
        remap = {0: 0}
        nxt = 1
        for k in order:
            if freq[k] >= 2:
                remap[k] = nxt
                nxt += 1
            else:
                remap[k] = 0
        expect = np.array([remap[x] for x in lab.tolist()], dtype=np.int64)
        for fn in (O.sort_by_size, O.sort_by_size_np):
            got = lab.copy()
            fn(got)
            assert got.tolist() == expect.tolist()
        got = lab.copy()
        C.sort_by_size(got)
        assert got.tolist() == expect.tolist()


# ------------------------------------------------------------------ end-to-end KATs

def test_readme_quickstart():
    q = K["quickstart"]
    for method in ("bfs", "components"):
        labels, _ = O.fit(q["points"], q["radius_cutoff"], q["cnn_cutoff"], method=method)
        assert labels.tolist() == q["labels"]


def test_basic_usage_records():
    b = K["basic_usage"]
    pts = K["quickstart"]["points"]
    labels, _ = O.fit(pts, 2.0, 1)
    assert labels.tolist() == b["labels_r2"]
    for rec in b["records"]:
        _, got = O.fit(pts, rec["radius_cutoff"], rec["cnn_cutoff"])
        assert got["n_clusters"] == rec["n_clusters"]
        assert round(got["ratio_largest"], 3) == rec["ratio_largest"]
        assert round(got["ratio_noise"], 3) == rec["ratio_noise"]


def test_sklearn_tutorial_records():
    data = sklearn_tutorial_datasets()
    for rec in K["sklearn_records"]["records"]:
        X = data[rec["dataset"]]
        labels = C.fit(X, rec["radius_cutoff"], rec["cnn_cutoff"], member_cutoff=rec["member_cutoff"])
        got = O.make_record(labels, len(X), rec["radius_cutoff"], rec["cnn_cutoff"], rec["member_cutoff"], None)
        assert got["n_clusters"] == rec["n_clusters"], rec["dataset"]
        assert f"{got['ratio_largest']:.3f}" == f"{rec['ratio_largest']:.3f}", rec["dataset"]
        assert f"{got['ratio_noise']:.3f}" == f"{rec['ratio_noise']:.3f}", rec["dataset"]


def test_moons_against_generator_truth():
    """test_procedures.py:127-174, moons case: max_diff 0 against the generator labels."""
    from sklearn import datasets
    from sklearn.preprocessing import StandardScaler

    X, truth = datasets.make_moons(n_samples=1500)
    X = StandardScaler().fit_transform(X)
    truth = truth + 1
    labels = C.fit(X, 0.2, 5)
    found = equalise_labels(labels)
    truth[found == 0] = 0
    assert (found == equalise_labels(truth)).all()


# ------------------------------------------------------------------ BFS == components

@pytest.mark.parametrize("seed", range(6))
def test_bfs_equals_components(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(50, 400))
    d = int(rng.integers(1, 5))
    X = np.concatenate([rng.normal(rng.uniform(-3, 3, d), rng.uniform(0.2, 1.0), size=(n // 2, d)),
                        rng.uniform(-4, 4, size=(n - n // 2, d))])
    r = float(rng.uniform(0.2, 1.2))
    c = int(rng.integers(0, 8))
    params = O.make_parameters(r, c, current_start=int(rng.integers(1, 5)))
    consider = (rng.uniform(size=n) > 0.1).astype(np.uint8)
    neigh = O.neighbourhoods(X, r)
    row_ptr, col = O.neighbourhoods_csr(X, r)
    l_bfs = np.zeros(n, dtype=np.int64)
    c_bfs = consider.copy()
    O.fit_bfs(neigh, params, l_bfs, c_bfs)
    l_cc = np.zeros(n, dtype=np.int64)
    c_cc = consider.copy()
    O.fit_components(row_ptr, col, params, l_cc, c_cc)
    assert (l_bfs == l_cc).all()
    assert (c_bfs == c_cc).all() and not c_cc.any()
    # C restatements (literal BFS and parallel components) agree as well
    rp32, col32 = row_ptr, col.astype(np.int32)
    for method in ("bfs", "cc"):
        lab, con = C.labels_from_csr(rp32, col32, params["similarity_cutoff"], params["n_member_cutoff"],
                                     consider=consider, current_start=params["current_start"], method=method)
        assert (lab == l_bfs).all(), method
        assert not con.any()


@pytest.mark.parametrize("d,use_grid", [(1, True), (2, True), (3, True), (3, False), (8, False)])
def test_c_oracle_neighbourhoods_match_numpy(d, use_grid):
    rng = np.random.default_rng(d)
    X = rng.normal(size=(700, d))
    X[5] = X[6]          # duplicate point
    X[7, 0] = X[8, 0] + 0.5
    rp, col = O.neighbourhoods_csr(X, 0.5)
    rp2, col2 = C.neighbours_csr(X, 0.5, use_grid=use_grid)
    assert (rp == rp2).all() and (col == col2).all()


@pytest.mark.parametrize("d", [1, 3, 7, 8, 9, 16, 64, 65])
def test_c_oracle_brute_force_early_exit_matches_numpy(d):
    """The SIMD brute force with its early exit (used at full workload sizes)
    gives the literal restatement's sets, including non-finite and huge
    coordinates, duplicates, pairs at exactly r and a row subset."""
    rng = np.random.default_rng(100 + d)
    X = rng.normal(size=(403, d))
    X[5] = X[6]
    X[9, 0] = np.nan
    X[11, d // 2] = np.inf
    X[13] = 1e200
    X[20] = X[21]
    X[20, 0] = X[21, 0] + 0.75           # exactly r apart along one axis
    X[22] = X[21]
    X[22, 0] = np.nextafter(X[21, 0] + 0.75, np.inf)
    r = 0.75 if d == 1 else 1.1 * np.sqrt(d)
    if d > 1:
        X[20, 1:] = X[21, 1:]
        X[22, 1:] = X[21, 1:]
        r = 0.75 if d < 9 else r
    rp, col = O.neighbourhoods_csr(X, r)
    rp2, col2 = C.neighbours_bf(X, r)
    assert (rp == rp2).all() and (col == col2).all()
    rows = rng.integers(0, len(X), size=37)
    rp3, col3 = C.neighbours_bf(X, r, rows=rows)
    for t, i in enumerate(rows):
        assert (col3[rp3[t]:rp3[t + 1]] == col[rp[i]:rp[i + 1]]).all()


# ------------------------------------------------------------------ predict

def test_predict_moons_reference_case():
    """test_procedures.py:246-277: every second point of a two-moons set,
    labelled with the generator's labels, predicts the generator's labels of
    the whole set at r=0.2, c=2."""
    from sklearn import datasets
    from sklearn.preprocessing import StandardScaler

    X, truth = datasets.make_moons(n_samples=1500, random_state=7)
    X = StandardScaler().fit_transform(X)
    truth = truth + 1
    lab_y, con_y = O.predict_firstmatch(X[::2], truth[::2], X, 0.2, 2)
    assert (equalise_labels(lab_y) == equalise_labels(truth)).all()
    assert (con_y == 0).all()


def _predict_toy():
    # reference points: 0 at (-0.95, 0) and 11 at (-1.9, 0) (label 3), 1-5 a tight
    # group at x = 0 (label 1), 6-10 a tight group at x = 0.3 (label 2); r = 1.
    # Neighbourhoods: N(0) = {0..5, 11}, N(1..5) = {0..10}, N(6..10) = {1..10},
    # N(11) = {0, 11}; y = (0.04, 0) sees {0..10}.
    X = np.array([[-0.95, 0.0]] + [[0.0, 0.01 * k] for k in range(5)]
                 + [[0.3, 0.01 * k] for k in range(5)] + [[-1.9, 0.0]])
    labels = np.array([3] + [1] * 5 + [2] * 5 + [3])
    return X, labels


def test_predict_first_match_rules():
    X, labels = _predict_toy()
    y = np.array([[0.04, 0.0], [0.04, 0.0], [5.0, 5.0]])
    # c = 3 (more than 4 neighbours, 5 shared): member 0 has 7 and shares 6 -> label 3
    lab, con = O.predict_firstmatch(X, labels, y[:1], 1.0, 3)
    assert lab.tolist() == [3] and con.tolist() == [0]
    # c = 5 (more than 6, 7 shared): member 0 has 7 but shares 6 -> member 1 gives 1
    lab, con = O.predict_firstmatch(X, labels, y[:1], 1.0, 5)
    assert lab.tolist() == [1]
    # c = 6 (more than 7): member 0 is skipped for its size -> member 1 gives 1
    lab, con = O.predict_firstmatch(X, labels, y[:1], 1.0, 6)
    assert lab.tolist() == [1]
    # only label 2 allowed: member 6, the first of the second group
    lab, con = O.predict_firstmatch(X, labels, y[:1], 1.0, 3, clusters={2})
    assert lab.tolist() == [2]
    # not considered: untouched; far point: no neighbours, untouched
    lab, con = O.predict_firstmatch(X, labels, y[1:], 1.0, 3, labels_y=[7, 0], consider_y=[0, 1])
    assert lab.tolist() == [7, 0] and con.tolist() == [0, 1]
    # c = 10: y needs more than 11 neighbours -> nothing assigned
    lab, con = O.predict_firstmatch(X, labels, y[:1], 1.0, 10)
    assert lab.tolist() == [0] and con.tolist() == [1]


# ------------------------------------------------------------------ other metrics

def _sqrt_threshold(r):
    from commonnnclustering_amd.engine import sqrt_decision_threshold
    return sqrt_decision_threshold(r)


def test_sqrt_threshold_exact():
    """fl(sqrt(s)) <= r  <=>  s <= T(r) on every double in a window around r*r."""
    import math
    rng = np.random.default_rng(0)
    radii = [0.0, 1.0, 0.1, 0.3, 2.5, 1e-160, 1e150, 3.0000000000000004] + list(rng.uniform(0, 10, 40))
    for r in radii:
        T = _sqrt_threshold(r)
        s = r * r
        for _ in range(300):
            s = math.nextafter(s, -math.inf)
        for _ in range(600):
            if s >= 0.0:  # sums of squares are never negative
                assert (math.sqrt(s) <= r) == (s <= T), (r, s, T)
            s = math.nextafter(s, math.inf)
    assert _sqrt_threshold(-1.0) < 0 and _sqrt_threshold(float("nan")) < 0
    assert _sqrt_threshold(float("inf")) == float("inf")


def test_sqrt_metric_equals_literal_loop():
    import math
    rng = np.random.default_rng(3)
    X = rng.normal(size=(120, 3))
    r = 0.9
    got = O.neighbourhoods_metric(X, r, "euclidean")
    for i in range(len(X)):
        want = [j for j in range(len(X)) if math.sqrt(O.reduced_distance_scalar(X[i], X[j])) <= r]
        assert got[i].tolist() == want


def test_sqrt_metric_boundary_differs_from_reduced():
    """Pairs with s just above fl(r*r) whose square root still rounds to r are
    neighbours under the sqrt metric only."""
    import math
    r = 0.7
    T = _sqrt_threshold(r)
    assert T > r * r
    # points on the circle of radius r around the origin: some rounded sums land in (r*r, T]
    found = 0
    for theta in np.linspace(0.01, 1.5, 4000):
        p = np.array([r * math.cos(theta), r * math.sin(theta)])
        s = O.reduced_distance_scalar(np.zeros(2), p)
        if r * r < s <= T:
            X = np.array([[0.0, 0.0], p])
            assert O.neighbourhoods_metric(X, r, "euclidean")[0].tolist() == [0, 1]
            assert O.neighbourhoods(X, r)[0].tolist() == [0]
            found += 1
    assert found > 10


def test_precomputed_metric_equals_reduced():
    rng = np.random.default_rng(4)
    X = rng.normal(size=(200, 2))
    r = 0.5
    D = O.reduced_distances(X, X)
    a = O.neighbourhoods_metric(D, r * r, "precomputed")
    b = O.neighbourhoods(X, r)
    assert all((x == y).all() for x, y in zip(a, b))


def test_periodic_metric_literal_loop_and_closed_form():
    import math
    bounds = [360.0, 360.0, 0.0]
    rng = np.random.default_rng(8)
    X = np.round(rng.uniform(-700, 700, size=(80, 3)), 1)
    D = O.periodic_reduced_distances(X, X, bounds)
    for i in range(0, 80, 7):
        for j in range(80):
            total = 0.0
            for k in range(3):
                dist = abs(X[i, k] - X[j, k])
                if bounds[k] > 0:
                    dist = math.fmod(dist, bounds[k])
                    if dist > bounds[k] / 2:
                        dist = bounds[k] - dist
                total = total + dist * dist
            assert D[i, j] == total
    # minimum image: 179 and -179 are 2 apart, 10 and 370 coincide; the third axis is not periodic
    P = np.array([[179.0, 0.0, 0.0], [-179.0, 0.0, 0.0], [10.0, 370.0, 5.0], [10.0, 10.0, 5.0]])
    Dp = O.periodic_reduced_distances(P, P, bounds)
    assert Dp[0, 1] == 4.0 and Dp[2, 3] == 0.0
    # whole-period shifts of integer-valued points leave every neighbourhood unchanged
    Y = np.round(rng.uniform(-180, 180, size=(150, 2)))
    shift = 360.0 * rng.integers(-2, 3, size=(150, 2))
    a = O.neighbourhoods_metric(Y, 25.0, "periodic", bounds=[360.0, 360.0])
    b = O.neighbourhoods_metric(Y + shift, 25.0, "periodic", bounds=[360.0, 360.0])
    assert all((x == y).all() for x, y in zip(a, b))
