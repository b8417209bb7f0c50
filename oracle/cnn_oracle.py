"""CPU oracle for the CommonNN ``Clustering.fit`` hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this module, and only as the
checker.  The product path (``commonnnclustering_amd``) never imports it and
fails loudly when its HIP library is missing.

This is a restatement, in numpy and plain Python, of the arithmetic and
control flow of cnnclustering 0.5.0 (janjoswig/CommonNNClustering) for the
default ``"coordinates"`` recipe:

  input      recipes.py:52-147 (``prepare_points_from_parts``; float64, C order, :144)
  metric     _types.pyx:2253-2267 (``MetricExtEuclideanReduced._calc_distance``)
             _types.pyx:2286-2287 (``_adjust_radius``: r**2)
  getter     _types.pyx:1692-1710 (``NeighboursGetterExtBruteForce._get``: s <= r2,
             ascending j, self included)
  enough     _types.pyx:1561-1564 (``n_points > member_cutoff``)
  checker    _types.pyx:2580-2610 (``SimilarityCheckerExtSwitchContains._check``)
             _types.pyx:2654-2706 (``SimilarityCheckerExtScreensorted._check``)
  params     _fit.pyx:129-161 (``make_parameters``: r -> r*r, c -> c+2, m = c+1)
  fitter     _fit.pyx:408-491 (``FitterExtBFS._fit``)
  sort       _types.pyx:212-286 (``Labels.sort_by_size``)
  wrapper    cluster.pyx:256-392 (``Clustering.fit``), report.py:91-191 (``Record``)
  metrics    _types.pyx:2105-2127 (``MetricExtPrecomputed``), :2170-2207 (``MetricExtEuclidean``),
             :2296-2317 (``MetricExtEuclideanPeriodicReduced._calc_distance``)
  predict    _fit.pyx:1241-1291 (``PredictorFirstmatch.predict``), _types.pyx:1712-1731
             (``_get_other``), cluster.pyx:560-653 (``Clustering.predict``)
  hierarchy  _fit.pyx:1014-1187 (``HierarchicalFitterRepeat``), _bundle.pyx:275-352
             (``isolate``), _fit.pyx:733-1011 (``HierarchicalFitterMSTPrim``),
             _bundle.pyx:355-395 (``check_children``), _types.pyx:2825-2851
             (``PriorityQueueMaxHeap``)

How each part is pinned (tests/test_oracle.py):
  * metric: bit-identical to a literal scalar Python loop (IEEE double, the
    same sub/mul/add sequence with no FMA); error against an exact rational
    evaluation stays within the textbook bound gamma_{2d} * d^2.
  * neighbourhoods: boundary pairs at exactly r and one ulp beyond; symmetry;
    self-inclusion; README/notebook known answers.
  * similarity: the known-answer cases of test/unit/test_check_similarity.py:39-48.
  * labelling: the literal BFS restatement (``fit_bfs``) and the
    connected-components restatement (``fit_components``) agree on random
    inputs; README.md:71-93 and basic_usage.ipynb:356 known answers; the
    scikit-learn tutorial records (scikit_learn_datasets.ipynb cell 10 output).
  * sort_by_size: test/unit/test_types.py:98-117 known answers; tie order
    checked against ``collections.Counter.most_common`` semantics directly.
  * other metrics: the sqrt metric's lists equal a literal scalar loop with
    math.sqrt (IEEE, correctly rounded) and include pairs whose decision differs
    from the reduced metric's at r*r; the precomputed metric's lists equal the
    reduced metric's when fed the reduced distance matrix with r*r; the
    periodic metric equals a literal scalar loop with math.fmod, gives the
    minimum-image closed form on hand-placed points and is invariant under
    shifting points by whole periods.
  * predict: the reference's own integration case
    (test/integration/test_procedures.py:246-277: half of a two-moons set
    labelled by the generator predicts the whole set's generator labels) and
    hand-built cases whose first match is fixed by construction.
  * hierarchy: ``isolate`` and ``reel`` against test/unit/test_clustering.py's
    regression reports and expected labels (the product's host code, which the
    Repeat restatement's ``_isolate_nodes`` must agree with); every Repeat
    node's labels equal a plain fit at that step restricted to the node's
    points; the Prim forest's total weight equals scipy's maximum spanning
    forest and, with distinct weights, its edge set; every MST hierarchy node
    (distinct weights) is a single-linkage cluster: the component holding its
    merge edge among the tree edges of weight >= its lambda.  Folding of tied
    weights, lifting of a single child's children (with its lambda) and
    pruning below member_cutoff in ``_check_children`` are pinned by three
    merge-loop runs worked out by hand from _fit.pyx:945-1011 and
    _bundle.pyx:354-398 (tests/test_hierarchy.py, MST_KATS).
"""

from __future__ import annotations

from collections import Counter
from itertools import count

import numpy as np

__all__ = [
    "prepare_points",
    "adjust_radius",
    "make_parameters",
    "reduced_distance_scalar",
    "reduced_distances",
    "neighbourhoods",
    "neighbourhoods_csr",
    "check_similarity",
    "similarity_count",
    "fit_bfs",
    "fit_components",
    "sort_by_size",
    "make_record",
    "fit",
]


# ---------------------------------------------------------------- input / params

def prepare_points(data) -> np.ndarray:
    """Coerce point data to a C-contiguous float64 (n, d) array.

    Follows recipes.py:52-147 for the 2D-sequence case: the result is
    ``np.asarray(np.vstack(data), order="C", dtype=float64)`` (:144), so float32
    input is upcast exactly.  Sequences of parts are stacked.
    """
    if isinstance(data, np.ndarray) and data.ndim == 2:
        return np.asarray(data, order="C", dtype=np.float64)
    if len(data) == 0:
        return np.zeros((1, 0), dtype=np.float64)  # recipes.py:106-109: np.array([[]])
    first = data[0]
    try:
        len(first[0])
        parts = [np.asarray(x) for x in data]   # sequence of 2D parts
    except TypeError:
        try:
            len(first)
            parts = [np.asarray(data)]          # 2D sequence
        except TypeError:
            parts = [np.array([data])]          # one point
    return np.asarray(np.vstack(parts), order="C", dtype=np.float64)


def adjust_radius(radius_cutoff: float) -> float:
    """r -> r*r in float64 (_types.pyx:2286-2287; GCC folds pow(r, 2) to r*r)."""
    r = float(radius_cutoff)
    return r * r


def make_parameters(radius_cutoff: float, cnn_cutoff: int, current_start: int = 1) -> dict:
    """Restates ``FitterExtInterface.make_parameters`` (_fit.pyx:129-161).

    The brute-force getter is self-counting, so the member cutoff becomes
    ``c + 1`` and the similarity cutoff ``c + 2``.
    """
    c = int(cnn_cutoff)
    return {
        "radius_cutoff": adjust_radius(radius_cutoff),
        "similarity_cutoff": c + 2,
        "similarity_cutoff_continuous": float(c + 2),
        "n_member_cutoff": c + 1,
        "current_start": int(current_start),
    }


# ---------------------------------------------------------------- metric

def reduced_distance_scalar(a, b) -> float:
    """Literal scalar restatement of _types.pyx:2253-2267.

    ``total = 0; for k: total = total + (a_k - b_k) * (a_k - b_k)`` in IEEE
    double, every operation separately rounded.
    """
    total = 0.0
    for x, y in zip(a, b):
        t = float(x) - float(y)
        total = total + t * t
    return total


def reduced_distances(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """All-pairs reduced Euclidean distances, bit-identical to the scalar loop.

    numpy's elementwise subtract, multiply and add are each correctly rounded
    and never contracted, so looping over the dimension in order reproduces the
    reference's summation exactly (_types.pyx:2262-2265).
    """
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    total = np.zeros((A.shape[0], B.shape[0]), dtype=np.float64)
    for k in range(A.shape[1]):
        diff = A[:, k, None] - B[None, :, k]
        total = total + diff * diff
    return total


# ---------------------------------------------------------------- neighbours

def neighbourhoods(X: np.ndarray, radius_cutoff: float, chunk: int = 1024) -> list:
    """N(i) = {j : s_ij <= r*r}, ascending j, i included (_types.pyx:1692-1710)."""
    X = np.asarray(X, dtype=np.float64)
    r2 = adjust_radius(radius_cutoff)
    out = []
    for s in range(0, X.shape[0], chunk):
        D = reduced_distances(X[s:s + chunk], X)
        for row in D:
            out.append(np.nonzero(row <= r2)[0].astype(np.int64))
    return out


def periodic_reduced_distances(A: np.ndarray, B: np.ndarray, bounds) -> np.ndarray:
    """MetricExtEuclideanPeriodicReduced._calc_distance (_types.pyx:2296-2317) for
    all pairs: per dimension d = |a - b|; for a positive bound, d = fmod(d, bound)
    (exact; Python's % on the non-negative d agrees) and, when d > bound / 2,
    d = bound - d; total = total + d * d in ascending dimension."""
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    total = np.zeros((A.shape[0], B.shape[0]), dtype=np.float64)
    for k in range(A.shape[1]):
        dist = np.abs(A[:, k, None] - B[None, :, k])
        bound = float(bounds[k])
        if bound > 0:
            dist = np.fmod(dist, bound)
            dist = np.where(dist > bound / 2, bound - dist, dist)
        total = total + dist * dist
    return total


def neighbourhoods_metric(X: np.ndarray, radius_cutoff: float, metric: str = "euclidean_r",
                          chunk: int = 1024, bounds=None) -> list:
    """Brute-force neighbourhoods under the reference's other metrics:

    ``euclidean``    MetricExtEuclidean (_types.pyx:2170-2207): sqrt of the
                     reduced sum (IEEE sqrt) compared with r as given;
    ``precomputed``  MetricExtPrecomputed (_types.pyx:2105-2127): X is the
                     (n, n) distance matrix, D[i, j] compared with r as given;
    ``euclidean_r``  the reduced metric (``neighbourhoods``).
    """
    if metric == "euclidean_r":
        return neighbourhoods(X, radius_cutoff, chunk)
    X = np.asarray(X, dtype=np.float64)
    r = float(radius_cutoff)
    out = []
    if metric == "precomputed":
        for row in X:
            out.append(np.nonzero(row <= r)[0].astype(np.int64))
        return out
    if metric == "periodic":
        r2 = adjust_radius(radius_cutoff)
        for s in range(0, X.shape[0], chunk):
            D = periodic_reduced_distances(X[s:s + chunk], X, bounds)
            for row in D:
                out.append(np.nonzero(row <= r2)[0].astype(np.int64))
        return out
    if metric != "euclidean":
        raise ValueError(f"unknown metric {metric!r}")
    for s in range(0, X.shape[0], chunk):
        D = np.sqrt(reduced_distances(X[s:s + chunk], X))
        for row in D:
            out.append(np.nonzero(row <= r)[0].astype(np.int64))
    return out


def neighbourhoods_csr(X: np.ndarray, radius_cutoff: float, chunk: int = 1024, metric: str = "euclidean_r",
                       bounds=None):
    """The same neighbourhoods as CSR: (row_ptr int64[n+1], col int64[nnz])."""
    lists = neighbourhoods_metric(X, radius_cutoff, metric, chunk, bounds)
    deg = np.array([len(x) for x in lists], dtype=np.int64)
    row_ptr = np.zeros(len(lists) + 1, dtype=np.int64)
    np.cumsum(deg, out=row_ptr[1:])
    col = np.concatenate(lists) if lists else np.zeros(0, dtype=np.int64)
    return row_ptr, col


# ---------------------------------------------------------------- similarity

def check_similarity(members_a, members_b, c: int) -> bool:
    """Restates ``SimilarityCheckerExtSwitchContains._check`` (_types.pyx:2580-2610).

    ``c == 0`` passes; otherwise iterate the shorter container and count
    members contained in the other one, passing as soon as the count reaches c.
    """
    if c == 0:
        return True
    a, b = list(members_a), list(members_b)
    if len(b) < len(a):
        a, b = b, a
    lookup = set(b)
    common = 0
    for m in a:
        if m in lookup:
            common += 1
            if common == c:
                return True
    return False


def similarity_count(members_a, members_b) -> int:
    """Full common-member count (``SimilarityCheckerExtSwitchContains._get``, :2612-2634)."""
    a, b = list(members_a), list(members_b)
    if len(b) < len(a):
        a, b = b, a
    lookup = set(b)
    return sum(1 for m in a if m in lookup)


# ---------------------------------------------------------------- fitters

def fit_bfs(neigh: list, params: dict, labels: np.ndarray, consider: np.ndarray) -> None:
    """Literal restatement of ``FitterExtBFS._fit`` (_fit.pyx:408-491).

    ``neigh[i]`` is N(i) as returned by the brute-force getter; the getter is
    deterministic, so recomputing it on every call (as the reference does) and
    looking it up give identical lists.  ``labels`` (int64) and ``consider``
    (uint8) are updated in place.  FIFO queue as ``QueueExtFIFOQueue``
    (_types.pyx:2802-2819).
    """
    m_cut = params["n_member_cutoff"]
    c_sim = params["similarity_cutoff"]
    current = params["current_start"]
    n = len(neigh)
    queue = []
    head = 0
    for init_point in range(n):
        if consider[init_point] == 0:
            continue
        consider[init_point] = 0
        na = neigh[init_point]
        if not len(na) > m_cut:
            continue
        labels[init_point] = current
        while True:
            for member in na:
                if consider[member] == 0:
                    continue
                nb = neigh[member]
                if not len(nb) > m_cut:
                    consider[member] = 0
                    continue
                if check_similarity(na, nb, c_sim):
                    consider[member] = 0
                    labels[member] = current
                    queue.append(member)
            if head == len(queue):
                break
            point = queue[head]
            head += 1
            na = neigh[point]
        current += 1


def fit_components(row_ptr: np.ndarray, col: np.ndarray, params: dict,
                   labels: np.ndarray, consider: np.ndarray) -> None:
    """Connected-components restatement of the same fit.

    Points that are considered and have more than ``n_member_cutoff``
    neighbours are core; an edge joins core points i, j when j is in N(i) and
    the similarity check passes.  Components are numbered from
    ``current_start`` in ascending order of their smallest point index, which
    is the order in which the BFS of _fit.pyx:421-491 opens clusters.  After
    the fit every ``consider`` flag is 0, as in the BFS (:422).
    """
    n = len(row_ptr) - 1
    m_cut = params["n_member_cutoff"]
    c_sim = params["similarity_cutoff"]
    deg = np.diff(row_ptr)
    core = (deg > m_cut) & (np.asarray(consider) != 0)
    parent = np.arange(n, dtype=np.int64)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for i in np.nonzero(core)[0]:
        ni = col[row_ptr[i]:row_ptr[i + 1]]
        for j in ni:
            if j <= i or not core[j]:
                continue
            ri, rj = find(i), find(j)
            if ri == rj:
                continue
            if check_similarity(ni, col[row_ptr[j]:row_ptr[j + 1]], c_sim):
                lo, hi = (ri, rj) if ri < rj else (rj, ri)
                parent[hi] = lo
    current = params["current_start"]
    label_of_root = {}
    for i in np.nonzero(core)[0]:   # ascending i: first visit of a root = its min index
        r = find(i)
        if r not in label_of_root:
            label_of_root[r] = current
            current += 1
        labels[i] = label_of_root[r]
    consider[:] = 0


# ---------------------------------------------------------------- sort_by_size

def sort_by_size(labels: np.ndarray, member_cutoff=None, max_clusters=None) -> None:
    """Restates ``Labels.sort_by_size`` (_types.pyx:212-286) in place.

    ``Counter(labels).most_common()`` sorts by count descending and keeps
    first-occurrence order among equal counts; label 0 is dropped first
    (:239-241).  Ranks beyond ``max_clusters`` and clusters smaller than
    ``member_cutoff`` (default 2, :235-237) become 0; survivors are
    renumbered 1, 2, ... (:254-268).
    """
    if member_cutoff is None:
        member_cutoff = 2
    frequencies = Counter(labels.tolist())
    frequencies.pop(0, None)
    if not frequencies:
        return
    if max_clusters is None:
        max_clusters = len(frequencies)
    reassign = {0: 0}
    new_labels = count(1)
    for rank, (old, size) in enumerate(frequencies.most_common(), 1):
        if rank > max_clusters:
            reassign[old] = 0
        elif size >= member_cutoff:
            reassign[old] = next(new_labels)
        else:
            reassign[old] = 0
    labels[:] = np.array([reassign[x] for x in labels.tolist()], dtype=labels.dtype)


def sort_by_size_np(labels: np.ndarray, member_cutoff=None, max_clusters=None) -> None:
    """Vectorised equivalent of :func:`sort_by_size` for large n.

    Key per cluster: (size descending, first occurrence ascending), which is
    exactly the order ``Counter.most_common`` yields (stable sort on a
    first-occurrence-ordered mapping).
    """
    if member_cutoff is None:
        member_cutoff = 2
    lab = np.asarray(labels)
    uniq, first, inv, cnt = np.unique(lab, return_index=True, return_inverse=True,
                                      return_counts=True)
    keep = uniq != 0
    if not keep.any():
        return
    if max_clusters is None:
        max_clusters = int(keep.sum())
    idx = np.nonzero(keep)[0]
    order = idx[np.lexsort((first[idx], -cnt[idx]))]
    new = np.zeros(len(uniq), dtype=np.int64)
    ranks = np.arange(1, len(order) + 1)
    ok = (ranks <= max_clusters) & (cnt[order] >= member_cutoff)
    new[order[ok]] = np.arange(1, int(ok.sum()) + 1)
    labels[:] = new[inv].astype(labels.dtype)


# ---------------------------------------------------------------- record / fit

def make_record(labels: np.ndarray, n_points: int, radius_cutoff, cnn_cutoff,
                member_cutoff, max_clusters, execution_time=None) -> dict:
    """The ``Record`` fields computed at cluster.pyx:366-390 (report.py:91-191)."""
    freq = Counter(np.asarray(labels).tolist())
    n_noise = freq.pop(0, 0)
    n_largest = freq.most_common(1)[0][1] if freq else 0
    return {
        "n_points": n_points,
        "radius_cutoff": radius_cutoff,
        "cnn_cutoff": cnn_cutoff,
        "member_cutoff": member_cutoff,
        "max_clusters": max_clusters,
        "n_clusters": len(set(np.asarray(labels).tolist()) - {0}),
        "ratio_largest": n_largest / n_points if n_points else 0.0,
        "ratio_noise": n_noise / n_points if n_points else 0.0,
        "execution_time": execution_time,
    }


def fit(data, radius_cutoff, cnn_cutoff, member_cutoff=None, max_clusters=None,
        cnn_offset=None, sort=True, method="components", metric="euclidean_r", bounds=None):
    """End-to-end restatement of ``Clustering.fit`` (cluster.pyx:256-392).

    Returns ``(labels int64[n], record dict)``.  ``method`` selects the
    literal BFS (small n) or the connected-components form.
    """
    X = prepare_points(data)
    if cnn_offset is None:
        cnn_offset = 0
    cnn_cutoff = cnn_cutoff - cnn_offset
    assert cnn_cutoff >= 0
    n = X.shape[0]
    labels = np.zeros(n, dtype=np.int64)
    consider = np.ones(n, dtype=np.uint8)
    params = make_parameters(radius_cutoff, cnn_cutoff, 1)
    if method == "bfs":
        fit_bfs(neighbourhoods_metric(X, radius_cutoff, metric, bounds=bounds), params, labels, consider)
    else:
        row_ptr, col = neighbourhoods_csr(X, radius_cutoff, metric=metric, bounds=bounds)
        fit_components(row_ptr, col, params, labels, consider)
    if sort:
        sort_by_size_np(labels, member_cutoff, max_clusters)
    rec = make_record(labels, n, radius_cutoff, cnn_cutoff, member_cutoff, max_clusters)
    return labels, rec


def neighbourhoods_other(X: np.ndarray, Y: np.ndarray, radius_cutoff: float, chunk: int = 1024) -> list:
    """For each y of Y: {i of X : s(y, i) <= r*r}, ascending i (_types.pyx:1712-1731).

    ``_calc_distance_other`` (_types.pyx:2269-2284) subtracts in the order
    (y_k - x_k); the square of a difference does not depend on the order.
    """
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    r2 = adjust_radius(radius_cutoff)
    out = []
    for s in range(0, Y.shape[0], chunk):
        D = reduced_distances(Y[s:s + chunk], X)
        for row in D:
            out.append(np.nonzero(row <= r2)[0].astype(np.int64))
    return out


def predict_firstmatch(X, labels_x, Y, radius_cutoff, cnn_cutoff, clusters=None, labels_y=None,
                       consider_y=None):
    """Restates ``PredictorFirstmatch.predict`` (_fit.pyx:1241-1291) with
    brute-force getters, as ``Clustering.predict`` calls it (cluster.pyx:600-630):
    parameters from make_parameters (r -> r*r, similarity c+2, members c+1),
    ``clusters`` defaults to the reference labels without 0.

    For each y with consider != 0: its reference neighbours must be more than
    n_member_cutoff; they are visited in ascending index, skipping labels not in
    the set and members whose own neighbourhood is not larger than
    n_member_cutoff; the first member passing the similarity check gives y its
    label and clears y's consider flag.  Returns (labels_y, consider_y).
    """
    X = prepare_points(X)
    Y = prepare_points(Y)
    labels_x = np.asarray(labels_x, dtype=np.int64)
    params = make_parameters(radius_cutoff, cnn_cutoff, 0)
    if clusters is None:
        clusters = set(int(v) for v in labels_x) - {0}
    clusters = set(int(v) for v in clusters)
    n_y = Y.shape[0]
    labels_y = np.zeros(n_y, dtype=np.int64) if labels_y is None else np.array(labels_y, dtype=np.int64)
    consider_y = np.ones(n_y, dtype=np.uint8) if consider_y is None else np.array(consider_y, dtype=np.uint8)
    m = params["n_member_cutoff"]
    c = params["similarity_cutoff"]
    if X.shape[0] == 0 or n_y == 0:
        return labels_y, consider_y
    own = neighbourhoods(X, radius_cutoff)
    other = neighbourhoods_other(X, Y, radius_cutoff)
    for point in range(n_y):
        if consider_y[point] == 0:
            continue
        na = other[point]
        if not len(na) > m:
            continue
        for member in na:
            label = int(labels_x[member])
            if label not in clusters:
                continue
            nb = own[member]
            if not len(nb) > m:
                continue
            if check_similarity(na, nb, c):
                consider_y[point] = 0
                labels_y[point] = label
                break
    return labels_y, consider_y


def gamma(m: int, u: float) -> float:
    """Higham's gamma_m = m u / (1 - m u)."""
    return m * u / (1.0 - m * u)


def exact_reduced_distance(a, b):
    """Exact rational value of sum_k (a_k - b_k)^2 as a Python Fraction."""
    from fractions import Fraction
    return sum((Fraction(float(x)) - Fraction(float(y))) ** 2 for x, y in zip(a, b))




# ---------------------------------------------------------------- hierarchy

_sort_labels = sort_by_size  # the Repeat restatement's keyword argument shadows the name
#
# Hierarchy nodes are plain dicts:
#   Repeat:  {"labels": list | None, "root": ndarray | None, "parent": ndarray | None,
#             "children": {label: node}}
#   MST:     {"nodes": set, "edges": {(u, v): weight} with u < v, "lambda": weight,
#             "children": {label: node}}

def _isolate_nodes(node: dict) -> None:
    """``isolate`` without input data (_bundle.pyx:297-319): one child per
    label except 0, in order of first occurrence, with parent indices into
    the node's points and root indices through the node's own."""
    node["children"] = {}
    mapping = {}
    for index, label in enumerate(node["labels"]):
        mapping.setdefault(int(label), []).append(index)
    for label, indices in mapping.items():
        if label == 0:
            continue
        parent = np.array(indices, dtype=np.intp)
        root = parent if node["root"] is None else node["root"][parent]
        node["children"][label] = {"labels": None, "root": root, "parent": parent, "children": {}}


def hierarchical_repeat(data, radius_cutoff, cnn_cutoff, member_cutoff=None, max_clusters=None,
                        cnn_offset=None, sort_by_size=True, bounds=None) -> dict:
    """Restates ``HierarchicalFitterRepeat._fit`` (_fit.pyx:1055-1187).

    ``HierarchicalFitter.make_parameters`` (_fit.pyx:95-126) reads
    ``self._neighbours_getter``, which the Repeat fitter does not have, so the
    radius and cnn_cutoff pass unchanged: N(i) = {j : s_ij <= radius_cutoff},
    similarity cutoff c and member cutoff c (no self-counting shift).  Each
    step is a literal BFS fit (``fit_bfs``) on fresh labels followed by
    ``sort_by_size``.  The parents of a step are visited in ascending label;
    the reference's std::unordered_map order is unspecified and matters only
    when a cluster has points under several parents.  With ``bounds`` the
    periodic reduced metric replaces the reduced one.  Returns the root node.
    """
    X = prepare_points(data)
    radii = [float(x) for x in (radius_cutoff if np.iterable(radius_cutoff) else [radius_cutoff])]
    cuts = [int(x) for x in (cnn_cutoff if np.iterable(cnn_cutoff) else [cnn_cutoff])]
    if len(radii) == 1:
        radii *= len(cuts)
    if len(cuts) == 1:
        cuts *= len(radii)
    assert len(radii) == len(cuts)
    offset = 0 if cnn_offset is None else cnn_offset
    n = X.shape[0]
    D = reduced_distances(X, X) if bounds is None else periodic_reduced_distances(X, X, bounds)
    root = {"labels": None, "root": None, "parent": None, "children": {}}
    previous = np.ones(n, dtype=np.int64)
    terminal = {1: root}
    for r, c in zip(radii, cuts):
        c_used = c - offset
        neigh = [np.nonzero(row <= r)[0] for row in D]
        params = {"similarity_cutoff": c_used, "n_member_cutoff": c_used, "current_start": 1}
        labels = np.zeros(n, dtype=np.int64)
        fit_bfs(neigh, params, labels, np.ones(n, dtype=np.uint8))
        if sort_by_size:
            _sort_labels(labels, member_cutoff, max_clusters)
        parent_map = {}
        for i in range(n):
            if previous[i] != 0:
                parent_map.setdefault(int(previous[i]), []).append(int(labels[i]))
        new_terminal = {}
        for p in sorted(parent_map):
            node = terminal[p]
            node["labels"] = parent_map[p]
            _isolate_nodes(node)
            new_terminal.update(node["children"])
        terminal = new_terminal
        previous = labels
    return root


def similarity_weights(neigh: list) -> list:
    """Per list entry b of row a: |N(a) ∩ N(b)| (``SimilarityChecker._get``,
    _types.pyx:2612-2634), the MST edge weight of _fit.pyx:851-856, 888-893."""
    return [[similarity_count(neigh[a], neigh[b]) for b in neigh[a]] for a in range(len(neigh))]


def mst_prim(neigh: list, weights: list = None) -> list:
    """Restates the Prim loop of ``HierarchicalFitterMSTPrim._fit``
    (_fit.pyx:832-915) with ``PriorityQueueMaxHeap`` (_types.pyx:2825-2851,
    heapq on (-weight, (a, b))).  Returns the tree edges (a, b, weight) in the
    order the tree queue pops them (_fit.pyx:937-939)."""
    import heapq

    if weights is None:
        weights = similarity_weights(neigh)
    n = len(neigh)
    consider = [1] * n
    queue, tree = [], []

    def push_members(point):
        for t, member in enumerate(neigh[point]):
            if consider[member] == 0:
                continue
            heapq.heappush(queue, (-weights[point][t], (point, int(member))))

    for point in range(n):
        if consider[point] == 0:
            continue
        consider[point] = 0
        push_members(point)
        while queue:
            weight, (a, b) = heapq.heappop(queue)
            if consider[b] == 0:
                continue
            consider[b] = 0
            heapq.heappush(tree, (weight, (a, b)))
            push_members(b)
    out = []
    while tree:
        weight, (a, b) = heapq.heappop(tree)
        out.append((a, b, -weight))
    return out


def _graph_size(node: dict) -> int:
    return len(node["nodes"])


def _check_children(node: dict, member_cutoff: int, needs_folding: bool) -> None:
    """Restates ``check_children`` (_bundle.pyx:355-395) on dict nodes."""
    from collections import deque

    if needs_folding:
        leafs, queue = [], deque()
        for child in node["children"].values():
            if child["lambda"] == node["lambda"]:
                queue.append(child)
            else:
                leafs.append(child)
        while queue:
            candidate = queue.popleft()
            for child in candidate["children"].values():
                if child["lambda"] == node["lambda"]:
                    queue.append(child)
                else:
                    leafs.append(child)
        node["children"] = {}
        for k, child in enumerate(leafs, 1):
            node["children"][k] = child
    kept = {}
    for k, v in enumerate(node["children"].values(), 1):
        if _graph_size(v) >= member_cutoff:
            kept[k] = v
    node["children"] = kept
    if len(node["children"]) == 1:
        _, child = node["children"].popitem()
        for label, grandchild in enumerate(child["children"].values(), 1):
            node["children"][label] = grandchild
            node["lambda"] = child["lambda"]


def mst_hierarchy(n: int, tree_edges: list, member_cutoff: int = 10) -> dict:
    """Restates the hierarchy construction of ``HierarchicalFitterMSTPrim._fit``
    (_fit.pyx:918-1011): merges the tree edges in order with a union-find
    (parent pointers, no compression, the union points the root of a at the
    root of b), one top node per edge holding the edge plus the graphs of the
    top nodes it joins; pruning by ``_check_children`` before each drop of the
    weight.  Returns the root node."""
    parent = list(range(n))
    label_of = list(range(n))
    tops = {}
    checked = [False] * max(n - 1, 0)

    def get_root(p):
        q = parent[p]
        while q != p:
            p = q
            q = parent[p]
        return p

    current = np.inf
    folding = False
    i = -1
    for i, (a, b, weight) in enumerate(tree_edges):
        ra, rb = get_root(a), get_root(b)
        la, lb = label_of[ra], label_of[rb]
        if weight < current:
            for index, top in tops.items():
                if checked[index]:
                    continue
                _check_children(top, member_cutoff, folding)
                checked[index] = True
            folding = False
        else:
            folding = True
        top = {"nodes": {a, b}, "edges": {(min(a, b), max(a, b)): weight}, "lambda": weight, "children": {}}
        tops[i] = top
        label = 1
        for lx in (la, lb):
            if lx >= n:
                sub = tops.pop(lx - n)
                top["nodes"] |= sub["nodes"]
                top["edges"].update(sub["edges"])
                top["children"][label] = sub
                label += 1
        parent[ra] = rb
        label_of[rb] = i + n
        current = weight
    for index, top in tops.items():
        if not checked[index]:
            _check_children(top, member_cutoff, folding)
    root = {"nodes": set(), "edges": {}, "lambda": -np.inf, "children": {}}
    if i < n - 2:
        count_ = 1
        for top in tops.values():
            if _graph_size(top) >= member_cutoff:
                # _fit.pyx:1001 adds the root's own (empty) edge set to itself:
                # a forest's root keeps an empty graph
                root["children"][count_] = top
                count_ += 1
    else:
        top = tops[n - 2]
        root["nodes"], root["edges"], root["lambda"] = top["nodes"], top["edges"], top["lambda"]
        root["top"] = top  # the reference does not attach the tree's hierarchy to the root
    return root
