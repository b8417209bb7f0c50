"""Device-level API: torch tensors on the GPU in, torch tensors out.

Thin wrappers over the C-ABI in ``include/cnn_gpu.h``.  Radii are passed
squared (``r2 = r * r`` in double), exactly the value the reference's
``make_parameters`` stores in ``ClusterParameters.radius_cutoff``
(_fit.pyx:136-141).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import GraphStats, LabelStats, get_context, ptr

__all__ = [
    "NeighbourGraph",
    "ScanGraph",
    "build_graph",
    "radius_csr",
    "labels_from_csr",
    "sort_by_size",
    "fit",
    "fit_batch",
    "fit_scan",
    "predict",
    "precomputed_csr",
    "sqrt_decision_threshold",
    "to_device_points",
]


def to_device_points(X, device=None) -> torch.Tensor:
    """Coordinates as a contiguous float32/float64 CUDA tensor (dtype kept)."""
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    if isinstance(X, torch.Tensor):
        t = X
        if t.dtype not in (torch.float32, torch.float64):
            t = t.to(torch.float64)
        return t.to(device).contiguous()
    arr = np.asarray(X)
    if arr.dtype not in (np.float32, np.float64):
        arr = arr.astype(np.float64)
    if arr.ndim == 1:
        arr = arr.reshape(1, -1)
    return torch.from_numpy(np.ascontiguousarray(arr)).to(device)


def _check_points(X: torch.Tensor):
    if not X.is_cuda:
        raise ValueError("X must be a CUDA (HIP) tensor")
    if X.dim() != 2:
        raise ValueError(f"X must be 2-D (n, d), got shape {tuple(X.shape)}")
    if not X.is_contiguous():
        raise ValueError("X must be contiguous (row-major)")
    return X.shape[0], X.shape[1], _lib.dtype_code(X)


class NeighbourGraph:
    """Resident radius graph on the device (cnn_graph handle).

    Neighbour sets are those of NeighboursGetterExtBruteForce._get
    (_types.pyx:1692-1710) with the reduced Euclidean metric at ``r2``.
    """

    def __init__(self, handle, ctx, X: torch.Tensor, r2: float, precision: int):
        self._handle = handle
        self._ctx = ctx
        self.n, self.d = X.shape
        self.device = X.device
        self.r2 = float(r2)
        self.precision = precision

    @property
    def handle(self):
        if self._handle is None:
            raise RuntimeError("graph was released")
        return self._handle

    @property
    def nnz(self) -> int:
        v = ctypes.c_int64()
        self._ctx.check(_lib.lib().cnn_graph_nnz(self.handle, ctypes.byref(v)))
        return v.value

    @property
    def stats(self) -> dict:
        s = GraphStats()
        self._ctx.check(_lib.lib().cnn_graph_stats_get(self.handle, ctypes.byref(s)))
        return s.to_dict()

    @property
    def nbytes(self) -> int:
        v = ctypes.c_int64()
        self._ctx.check(_lib.lib().cnn_graph_bytes(self.handle, ctypes.byref(v)))
        return v.value

    def to_csr(self):
        """(row_ptr int64[n+1], col int32[nnz]) in caller point order, rows ascending."""
        ctx = get_context(self.device)
        nnz = self.nnz
        row_ptr = torch.empty(self.n + 1, dtype=torch.int64, device=self.device)
        col = torch.empty(max(nnz, 1), dtype=torch.int32, device=self.device)
        ctx.check(_lib.lib().cnn_graph_export_csr(ctx.handle, self.handle, ptr(row_ptr), ptr(col), nnz),
                  "cnn_graph_export_csr")
        return row_ptr, col[:nnz]

    def rethreshold(self, r2: float, precision=None) -> "NeighbourGraph":
        ctx = get_context(self.device)
        p = self.precision if precision is None else _lib.precision_code(precision)
        h = ctypes.c_void_p()
        ctx.check(_lib.lib().cnn_graph_rethreshold(ctx.handle, self.handle, float(r2), p, ctypes.byref(h)),
                  "cnn_graph_rethreshold")
        g = NeighbourGraph.__new__(NeighbourGraph)
        g._handle, g._ctx = h, ctx
        g.n, g.d, g.device, g.r2, g.precision = self.n, self.d, self.device, float(r2), p
        return g

    def labels(self, c_sim: int, m_core: int, consider=None, current_start: int = 1, labels=None,
               sim_mode: int = _lib.CNN_SIM_AUTO):
        """In-place labelling (FitterExtBFS._fit semantics); returns (labels, stats)."""
        ctx = get_context(self.device)
        if labels is None:
            labels = torch.zeros(self.n, dtype=torch.int64, device=self.device)
        _check_labels(labels, self.n, consider)
        st = LabelStats()
        ctx.check(_lib.lib().cnn_graph_labels(ctx.handle, self.handle, int(c_sim), int(m_core), ptr(consider),
                                              int(current_start), int(sim_mode), ptr(labels), ctypes.byref(st)),
                  "cnn_graph_labels")
        return labels, st.to_dict()

    def fit(self, c_sim: int, m_core: int, sort: bool = True, member_cutoff=None, max_clusters=None,
            labels=None):
        """Fresh labels for all points (optionally sort_by_size); returns (labels, stats)."""
        ctx = get_context(self.device)
        if labels is None:
            labels = torch.empty(self.n, dtype=torch.int64, device=self.device)
        _check_labels(labels, self.n, None)
        st = LabelStats()
        ctx.check(_lib.lib().cnn_graph_fit(ctx.handle, self.handle, int(c_sim), int(m_core), int(bool(sort)),
                                           _member_cutoff(member_cutoff), _max_clusters(max_clusters),
                                           ptr(labels), ctypes.byref(st)), "cnn_graph_fit")
        return labels, st.to_dict()

    def scan(self, radii, cnn_cutoffs, member_cutoff=None, max_clusters=None, labels=None) -> "ScanGraph":
        """Records the (r, c) scan of this graph into a hipGraph (cnn_graph_scan_create):
        radii (each <= the graph's radius) and cnn cutoffs c; see ScanGraph."""
        return ScanGraph(self, radii, cnn_cutoffs, member_cutoff, max_clusters, labels)

    def release(self) -> None:
        if self._handle is not None:
            _lib.lib().cnn_graph_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class ScanGraph:
    """A parameter scan of a resident graph recorded once into a hipGraph.

    ``launch()`` replays it: labels[ri, ci] are the labels of a fresh fit at
    radius radii[ri] and cnn_cutoff cnn_cutoffs[ci] (c_sim = c + 2,
    m_core = c + 1), sorted by size -- each radius's row words re-thresholded
    exactly in FP64, every labelling kernel launched once per radius for all
    cutoffs.  The graph must outlive the scan.  Needs a graph that keeps its
    row words (CNNCapacityError otherwise)."""

    def __init__(self, graph: "NeighbourGraph", radii, cnn_cutoffs, member_cutoff=None, max_clusters=None,
                 labels=None):
        self.graph = graph
        ctx = get_context(graph.device)
        self._ctx = ctx
        r2 = np.ascontiguousarray([float(r) * float(r) for r in radii], dtype=np.float64)
        cs = np.asarray(list(cnn_cutoffs), dtype=np.int64)
        c_sim = np.ascontiguousarray(cs + 2, dtype=np.int32)
        m_core = np.ascontiguousarray(cs + 1, dtype=np.int32)
        self.shape = (len(r2), len(cs), graph.n)
        if labels is None:
            labels = torch.zeros(self.shape, dtype=torch.int64, device=graph.device)
        elif tuple(labels.shape) != self.shape or labels.dtype != torch.int64 or not labels.is_contiguous():
            raise ValueError(f"labels must be a contiguous int64 tensor of shape {self.shape}")
        self.labels = labels
        h = ctypes.c_void_p()
        ctx.check(_lib.lib().cnn_graph_scan_create(ctx.handle, graph.handle, r2.ctypes.data, len(r2),
                                                   c_sim.ctypes.data, m_core.ctypes.data, len(cs),
                                                   _member_cutoff(member_cutoff), _max_clusters(max_clusters),
                                                   ptr(labels), ctypes.byref(h)), "cnn_graph_scan_create")
        self._handle = h

    def launch(self):
        """Replays the scan; returns (labels [n_r, n_c, n], label stats per pair)."""
        if self._handle is None:
            raise RuntimeError("scan was released")
        n_pairs = self.shape[0] * self.shape[1]
        ls = (LabelStats * n_pairs)()
        self._ctx.check(_lib.lib().cnn_scan_launch(self._ctx.handle, self._handle, ctypes.cast(ls, ctypes.c_void_p)),
                        "cnn_scan_launch")
        return self.labels, [ls[i].to_dict() for i in range(n_pairs)]

    def release(self) -> None:
        if self._handle is not None:
            _lib.lib().cnn_scan_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _member_cutoff(v) -> int:
    return 2 if v is None else int(v)


def _max_clusters(v) -> int:
    return -1 if v is None else int(v)


def _check_labels(labels, n, consider):
    if labels.dtype != torch.int64 or not labels.is_cuda or labels.numel() != n or not labels.is_contiguous():
        raise ValueError("labels must be a contiguous int64 CUDA tensor of length n")
    if consider is not None:
        if consider.dtype != torch.uint8 or not consider.is_cuda or consider.numel() != n:
            raise ValueError("consider must be a uint8 CUDA tensor of length n")


def build_graph(X: torch.Tensor, r2: float, precision="f64", bounds=None) -> NeighbourGraph:
    """Radius graph of X at squared radius r2.  ``bounds`` (one period per
    dimension, <= 0 for a non-periodic one) selects the periodic reduced
    metric, which has the plain FP64 form only."""
    n, d, dt = _check_points(X)
    p = _lib.precision_code(precision)
    ctx = get_context(X.device)
    h = ctypes.c_void_p()
    if bounds is None:
        ctx.check(_lib.lib().cnn_graph_build(ctx.handle, ptr(X), n, d, dt, float(r2), p, ctypes.byref(h)),
                  "cnn_graph_build")
    else:
        b = np.ascontiguousarray(bounds, dtype=np.float64).reshape(-1)
        if b.size != d:
            raise ValueError(f"bounds needs one entry per dimension ({d}), got {b.size}")
        p = _lib.precision_code("f64")
        ctx.check(_lib.lib().cnn_graph_build_periodic(ctx.handle, ptr(X), n, d, dt, float(r2), b.ctypes.data,
                                                      ctypes.byref(h)), "cnn_graph_build_periodic")
    return NeighbourGraph(h, ctx, X, r2, p)


def radius_csr(X: torch.Tensor, r2: float, precision="f64"):
    """Neighbour lists as CSR in caller order: (row_ptr int64, col int32, stats)."""
    g = build_graph(X, r2, precision)
    try:
        row_ptr, col = g.to_csr()
        return row_ptr, col, g.stats
    finally:
        g.release()


def labels_from_csr(row_ptr: torch.Tensor, col: torch.Tensor, c_sim: int, m_core: int, consider=None,
                    current_start: int = 1, labels=None, sim_mode: int = _lib.CNN_SIM_AUTO):
    """Labels from a caller CSR (rows sorted ascending, unique); in-place semantics."""
    n = row_ptr.numel() - 1
    device = row_ptr.device
    ctx = get_context(device)
    if row_ptr.dtype != torch.int64 or col.dtype != torch.int32:
        raise ValueError("row_ptr must be int64 and col int32")
    if not (row_ptr.is_cuda and col.is_cuda and row_ptr.is_contiguous() and col.is_contiguous()):
        raise ValueError("row_ptr and col must be contiguous CUDA tensors")
    if n > 0:
        if int(row_ptr[0]) != 0 or bool((row_ptr[1:] < row_ptr[:-1]).any()):
            raise ValueError("row_ptr must start at 0 and be non-decreasing")
        if int(row_ptr[-1]) > col.numel():
            raise ValueError(f"row_ptr[-1] = {int(row_ptr[-1])} exceeds len(col) = {col.numel()}")
    if labels is None:
        labels = torch.zeros(n, dtype=torch.int64, device=device)
    _check_labels(labels, n, consider)
    col_arg = col if col.numel() else torch.zeros(1, dtype=torch.int32, device=device)
    st = LabelStats()
    ctx.check(_lib.lib().cnn_labels_from_csr(ctx.handle, ptr(row_ptr), ptr(col_arg), n, int(c_sim), int(m_core),
                                             ptr(consider), int(current_start), int(sim_mode), ptr(labels),
                                             ctypes.byref(st)), "cnn_labels_from_csr")
    return labels, st.to_dict()


def sort_by_size(labels: torch.Tensor, member_cutoff=None, max_clusters=None) -> int:
    """Labels.sort_by_size (_types.pyx:212-286) in place on a CUDA int64 tensor."""
    if labels.dtype != torch.int64 or not labels.is_cuda or not labels.is_contiguous():
        raise ValueError("labels must be a contiguous int64 CUDA tensor")
    ctx = get_context(labels.device)
    k = ctypes.c_int64()
    ctx.check(_lib.lib().cnn_sort_by_size(ctx.handle, ptr(labels), labels.numel(), _member_cutoff(member_cutoff),
                                          _max_clusters(max_clusters), ctypes.byref(k)), "cnn_sort_by_size")
    return k.value


def fit(X: torch.Tensor, r2: float, c_sim: int, m_core: int, precision="f64", sort: bool = True,
        member_cutoff=None, max_clusters=None, labels=None, keep_graph: bool = False):
    """Fused fresh fit: returns (labels int64[n], graph or None, graph stats, label stats)."""
    n, d, dt = _check_points(X)
    p = _lib.precision_code(precision)
    ctx = get_context(X.device)
    if labels is None:
        labels = torch.empty(n, dtype=torch.int64, device=X.device)
    _check_labels(labels, n, None)
    gs, ls = GraphStats(), LabelStats()
    h = ctypes.c_void_p()
    ctx.check(_lib.lib().cnn_fit(ctx.handle, ptr(X), n, d, dt, float(r2), int(c_sim), int(m_core), p,
                                 int(bool(sort)), _member_cutoff(member_cutoff), _max_clusters(max_clusters),
                                 ptr(labels), ctypes.byref(h) if keep_graph else None,
                                 ctypes.byref(gs), ctypes.byref(ls)), "cnn_fit")
    graph = NeighbourGraph(h, ctx, X, r2, p) if keep_graph else None
    return labels, graph, gs.to_dict(), ls.to_dict()


def fit_batch(X: torch.Tensor, offsets, r2, c_sim, m_core, precision="f64", member_cutoff=None,
              max_clusters=None, labels=None):
    """Many datasets stacked row-wise in X; dataset b is rows offsets[b]:offsets[b+1]."""
    n, d, dt = _check_points(X)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    n_sets = len(off) - 1
    if n_sets < 0 or off[-1] != n:
        raise ValueError("offsets must start a partition of the rows of X")
    r2a = np.ascontiguousarray(np.broadcast_to(np.asarray(r2, dtype=np.float64), (n_sets,)))
    ca = np.ascontiguousarray(np.broadcast_to(np.asarray(c_sim, dtype=np.int32), (n_sets,)))
    ma = np.ascontiguousarray(np.broadcast_to(np.asarray(m_core, dtype=np.int32), (n_sets,)))
    p = _lib.precision_code(precision)
    ctx = get_context(X.device)
    if labels is None:
        labels = torch.empty(n, dtype=torch.int64, device=X.device)
    gs, ls = GraphStats(), LabelStats()
    ctx.check(_lib.lib().cnn_fit_batch(ctx.handle, ptr(X), off.ctypes.data, n_sets, d, dt, r2a.ctypes.data,
                                       ca.ctypes.data, ma.ctypes.data, p, _member_cutoff(member_cutoff),
                                       _max_clusters(max_clusters), ptr(labels), ctypes.byref(gs),
                                       ctypes.byref(ls)), "cnn_fit_batch")
    return labels, gs.to_dict(), ls.to_dict()


def fit_scan(X: torch.Tensor, radii, cnn_cutoffs, precision="f32_band", member_cutoff=None, max_clusters=None,
             labels=None):
    """Fits for every (radius_cutoff, cnn_cutoff) pair of one dataset.

    radii are radius cutoffs r (squared here, r2 = r*r as make_parameters,
    _fit.pyx:143); cnn_cutoffs are c (c_sim = c + 2, m_core = c + 1).  The
    graph is built once at max(r); returns (labels [n_r, n_c, n] int64 on the
    device, graph stats at max(r), list of label stats per pair).
    """
    n, d, dt = _check_points(X)
    r2 = np.ascontiguousarray([float(r) * float(r) for r in radii], dtype=np.float64)
    cs = np.asarray(list(cnn_cutoffs), dtype=np.int64)
    c_sim = np.ascontiguousarray(cs + 2, dtype=np.int32)
    m_core = np.ascontiguousarray(cs + 1, dtype=np.int32)
    n_r, n_c = len(r2), len(cs)
    if labels is None:
        labels = torch.empty((n_r, n_c, n), dtype=torch.int64, device=X.device)
    elif tuple(labels.shape) != (n_r, n_c, n) or labels.dtype != torch.int64 or not labels.is_contiguous():
        raise ValueError(f"labels must be a contiguous int64 tensor of shape {(n_r, n_c, n)}")
    p = _lib.precision_code(precision)
    ctx = get_context(X.device)
    gs = GraphStats()
    ls = (LabelStats * max(n_r * n_c, 1))()
    ctx.check(_lib.lib().cnn_fit_scan(ctx.handle, ptr(X), n, d, dt, r2.ctypes.data, n_r, c_sim.ctypes.data,
                                      m_core.ctypes.data, n_c, p, _member_cutoff(member_cutoff),
                                      _max_clusters(max_clusters), ptr(labels), ctypes.byref(gs),
                                      ctypes.cast(ls, ctypes.c_void_p)), "cnn_fit_scan")
    return labels, gs.to_dict(), [ls[i].to_dict() for i in range(n_r * n_c)]


def predict(X: torch.Tensor, labels_x: torch.Tensor, Y: torch.Tensor, r2: float, c_sim: int, m_core: int,
            clusters, labels_y: torch.Tensor = None, consider_y: torch.Tensor = None, precision="f64"):
    """Labels for the points Y from the labelled reference points X, as
    PredictorFirstmatch.predict (_fit.pyx:1241-1291): for every considered y,
    the first reference neighbour (ascending index) whose label is in
    ``clusters``, which has more than ``m_core`` reference neighbours and which
    shares at least ``c_sim`` of y's reference neighbours gives y its label and
    clears y's consider flag.  ``labels_y`` / ``consider_y`` (int64 / uint8 on
    the device) are updated in place (zeros / ones when None).  Returns
    (labels_y, consider_y, similarity tests executed).
    """
    n_x, d, dt = _check_points(X)
    n_y, d_y, dt_y = _check_points(Y)
    if d_y != d:
        raise ValueError("X and Y differ in dimension")
    if dt_y != dt:  # the reference works in float64: upcasting float32 is exact
        X, Y = X.double().contiguous(), Y.double().contiguous()
        n_x, d, dt = _check_points(X)
    if labels_x.dtype != torch.int64 or not labels_x.is_cuda or labels_x.numel() != n_x:
        raise ValueError("labels_x must be an int64 CUDA tensor of length n_x")
    labels_x = labels_x.contiguous()
    if labels_y is None:
        labels_y = torch.zeros(n_y, dtype=torch.int64, device=Y.device)
    if consider_y is None:
        consider_y = torch.ones(n_y, dtype=torch.uint8, device=Y.device)
    _check_labels(labels_y, n_y, consider_y)
    consider_y = consider_y.contiguous()
    wanted = sorted(int(c) for c in clusters if int(c) >= 0)
    n_allowed = (wanted[-1] + 1) if wanted else 0
    table = np.zeros(max(n_allowed, 1), dtype=np.uint8)
    table[wanted] = 1
    allowed = torch.from_numpy(table).to(X.device)
    ctx = get_context(X.device)
    tests = ctypes.c_int64(0)
    ctx.check(_lib.lib().cnn_predict(ctx.handle, ptr(X), n_x, ptr(Y), n_y, d, dt, float(r2),
                                     _lib.precision_code(precision), int(c_sim), int(m_core), ptr(labels_x),
                                     ptr(allowed), n_allowed, ptr(labels_y), ptr(consider_y),
                                     ctypes.byref(tests)), "cnn_predict")
    return labels_y, consider_y, int(tests.value)


def precomputed_csr(D: torch.Tensor, r: float):
    """Neighbour lists from a precomputed (n, n) distance matrix on the device:
    N(i) = {j : D[i, j] <= r} (MetricExtPrecomputed under the brute-force
    getter, _types.pyx:1692-1710, 2105-2127).  Returns (row_ptr int64[n+1],
    col int32[nnz], number of pairs {i, j} whose two decisions differ)."""
    n, n2, dt = _check_points(D)
    if n != n2:
        raise ValueError(f"a distance matrix must be square, got {tuple(D.shape)}")
    ctx = get_context(D.device)
    row_ptr = torch.empty(n + 1, dtype=torch.int64, device=D.device)
    nnz, asym = ctypes.c_int64(0), ctypes.c_int64(0)
    L = _lib.lib()
    rc = L.cnn_precomputed_csr(ctx.handle, ptr(D), n, dt, float(r), ptr(row_ptr), None, 0,
                               ctypes.byref(nnz), ctypes.byref(asym))
    if rc == _lib.CNN_ERR_CAPACITY and nnz.value > 0:
        col = torch.empty(nnz.value, dtype=torch.int32, device=D.device)
        rc = L.cnn_precomputed_csr(ctx.handle, ptr(D), n, dt, float(r), ptr(row_ptr), ptr(col), nnz.value,
                                   ctypes.byref(nnz), ctypes.byref(asym))
    else:
        col = torch.empty(0, dtype=torch.int32, device=D.device)
    ctx.check(rc, "cnn_precomputed_csr")
    return row_ptr, col, int(asym.value)


def _check_csr(row_ptr: torch.Tensor, col: torch.Tensor):
    if not (row_ptr.is_cuda and col.is_cuda):
        raise ValueError("row_ptr and col must be CUDA tensors")
    if row_ptr.dtype != torch.int64 or col.dtype != torch.int32:
        raise ValueError("row_ptr must be int64 and col int32")
    if not (row_ptr.is_contiguous() and col.is_contiguous()) or row_ptr.dim() != 1 or col.dim() != 1:
        raise ValueError("row_ptr and col must be contiguous 1-D tensors")
    n = row_ptr.shape[0] - 1
    if n < 0:
        raise ValueError("row_ptr must hold n + 1 offsets")
    return n


def edge_similarity(row_ptr: torch.Tensor, col: torch.Tensor) -> torch.Tensor:
    """weight[t] = |N(a) ∩ N(b)| for the entry t of row a holding b
    (SimilarityChecker._get, _types.pyx:2612-2634): int32[nnz] on the device.
    Rows must be strictly ascending (checked on the device)."""
    n = _check_csr(row_ptr, col)
    ctx = get_context(row_ptr.device)
    weight = torch.empty(max(col.shape[0], 1), dtype=torch.int32, device=col.device)
    if n > 0 and int(row_ptr[-1].item()) > col.shape[0]:
        raise ValueError("col holds fewer entries than row_ptr[n]")
    ctx.check(_lib.lib().cnn_csr_edge_similarity(ctx.handle, ptr(row_ptr), ptr(col), n, ptr(weight)),
              "cnn_csr_edge_similarity")
    return weight[:col.shape[0]]


def similarity_mst(row_ptr: torch.Tensor, col: torch.Tensor):
    """Maximum spanning forest of the similarity-weighted lists in the order
    of HierarchicalFitterMSTPrim's Prim loop (_fit.pyx:832-915): returns host
    int64 arrays (a, b, weight), sorted by (-weight, a, b)."""
    n = _check_csr(row_ptr, col)
    if n > 0 and int(row_ptr[-1].item()) > col.shape[0]:
        raise ValueError("col holds fewer entries than row_ptr[n]")
    ctx = get_context(row_ptr.device)
    cap = max(n - 1, 0)
    a = np.empty(max(cap, 1), dtype=np.int64)
    b = np.empty_like(a)
    w = np.empty_like(a)
    k = ctypes.c_int64(0)
    ctx.check(_lib.lib().cnn_similarity_mst(ctx.handle, ptr(row_ptr), ptr(col), n, a.ctypes.data, b.ctypes.data,
                                            w.ctypes.data, cap, ctypes.byref(k)), "cnn_similarity_mst")
    m = k.value
    return a[:m], b[:m], w[:m]


def sqrt_decision_threshold(r: float) -> float:
    """Largest double s with fl(sqrt(s)) <= r, for IEEE (correctly rounded)
    sqrt.  The non-reduced Euclidean metric decides sqrt(s_ij) <= r
    (MetricExtEuclidean, _types.pyx:2170-2207); since fl(sqrt(.)) is monotone,
    that holds exactly when s_ij <= this threshold, so the reduced-metric
    kernels (which evaluate s_ij in the reference's order) decide it exactly.
    Negative or NaN r: -1 (no sum qualifies)."""
    import math

    r = float(r)
    if r != r or r < 0.0:
        return -1.0
    if r == math.inf:
        return math.inf
    s = r * r
    while s > 0.0 and math.sqrt(s) > r:
        s = math.nextafter(s, -math.inf)
    while True:
        t = math.nextafter(s, math.inf)
        if t != math.inf and math.sqrt(t) <= r:
            s = t
        else:
            return s
