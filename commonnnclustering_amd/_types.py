"""Clustering components: parameters, labels, input data, metric, getters, checkers.

Mirrors the component interfaces of cnnclustering 0.5.0 ``_types.pyx`` so GPU
components take the place of the Cython ones:

  ClusterParameters        _types.pyx:22-85
  Labels (sort_by_size)    _types.pyx:88-286
  InputData*               _types.pyx:305-436, 1034-1109, 1153-1217
  Metric*                  _types.pyx:654-716, 2250-2287
  DistanceGetter*          _types.pyx:591-652, 1966-2017
  NeighboursGetter*        _types.pyx:478-589, 1658-1735, 1791-1832
  Neighbours*              _types.pyx:438-544, 1377-1424, 1525-1576
  SimilarityChecker*       _types.pyx:718-780, 2496-2753

The GPU components work on whole phases (all neighbourhoods at once, all edges
at once) rather than per point; the per-point accessors exist for inspection
and for code written against the reference's Python-level API.
"""

from __future__ import annotations

import heapq
from abc import ABC, abstractmethod
from collections import defaultdict
from typing import Optional, Sequence

import numpy as np

from ._primitive_types import P_ABOOL, P_AINDEX, P_AVALUE
from . import _lib


# ---------------------------------------------------------------- parameters

class ClusterParameters:
    """Input parameters of the clustering procedure (_types.pyx:22-85).

    ``radius_cutoff`` holds the metric-adjusted radius (r*r for the reduced
    Euclidean metric); ``similarity_cutoff`` and ``n_member_cutoff`` hold the
    self-counting-adjusted cutoffs produced by ``Fitter.make_parameters``.
    """

    def __init__(self, radius_cutoff: float, similarity_cutoff: int = 0,
                 similarity_cutoff_continuous: float = 0.0, n_member_cutoff: int = None,
                 current_start: int = 1):
        if n_member_cutoff is None:
            n_member_cutoff = similarity_cutoff
        self.radius_cutoff = float(radius_cutoff)
        self.similarity_cutoff = int(similarity_cutoff)
        self.similarity_cutoff_continuous = float(similarity_cutoff_continuous)
        self.n_member_cutoff = int(n_member_cutoff)
        self.current_start = int(current_start)

    def to_dict(self):
        return {
            "radius_cutoff": self.radius_cutoff,
            "similarity_cutoff": self.similarity_cutoff,
            "similarity_cutoff_continuous": self.similarity_cutoff_continuous,
            "n_member_cutoff": self.n_member_cutoff,
            "current_start": self.current_start,
        }

    def __repr__(self):
        return f"{self.to_dict()!r}"

    def __str__(self):
        return f"{self.to_dict()!s}"


# ---------------------------------------------------------------- labels

def label_counts(labels: np.ndarray):
    """(distinct labels ascending, their counts), by one bincount when the
    labels span fewer than 2^22 values (a sort otherwise)."""
    lab = np.asarray(labels)
    if lab.size == 0:
        return np.zeros(0, dtype=P_AINDEX), np.zeros(0, dtype=np.int64)
    lo, hi = int(lab.min()), int(lab.max())
    if hi - lo < (1 << 22):
        cnt = np.bincount(lab - lo)
        present = np.flatnonzero(cnt)
        return (present + lo).astype(P_AINDEX), cnt[present]
    uniq, cnt = np.unique(lab, return_counts=True)
    return uniq.astype(P_AINDEX), cnt


class Labels:
    """Cluster label assignments (_types.pyx:88-286).

    ``labels`` is int64, ``consider`` uint8.  ``sort_by_size`` runs on the GPU
    (cnn_sort_by_size); the host arrays are the source of truth between calls.
    """

    def __init__(self, labels, *, consider=None, meta=None):
        self._labels = np.ascontiguousarray(labels, dtype=P_AINDEX) if not (
            isinstance(labels, np.ndarray) and labels.dtype == P_AINDEX and labels.flags.c_contiguous) else labels
        if consider is None:
            self._consider = np.ones_like(self._labels, dtype=P_ABOOL)
        else:
            self._consider = np.ascontiguousarray(consider, dtype=P_ABOOL)
            if self._labels.shape[0] != self._consider.shape[0]:
                raise ValueError("'labels' and 'consider' must have the same length")
        self.meta = {} if meta is None else meta
        self._consider_set = set()

    @property
    def mapping(self):
        return self.to_mapping()

    @property
    def set(self):
        return self.to_set()

    @property
    def labels(self):
        return np.asarray(self._labels)

    @property
    def consider(self):
        return np.asarray(self._consider)

    @property
    def consider_set(self):
        return self._consider_set

    @consider_set.setter
    def consider_set(self, value):
        self._consider_set = value

    @property
    def n_points(self):
        return self._labels.shape[0]

    def __repr__(self):
        return f"{type(self).__name__}({list(self.labels)!s})"

    def __str__(self):
        return f"{self.labels!s}"

    @classmethod
    def from_sequence(cls, labels, *, consider=None, meta=None) -> "Labels":
        labels = np.array(labels, order="C", dtype=P_AINDEX)
        if consider is not None:
            consider = np.array(consider, order="C", dtype=P_ABOOL)
        return cls(labels, consider=consider, meta=meta)

    @classmethod
    def from_length(cls, n: int, meta=None) -> "Labels":
        return cls(np.zeros(n, order="C", dtype=P_AINDEX), meta=meta)

    def to_mapping(self):
        mapping = defaultdict(list)
        for index, label in enumerate(self._labels.tolist()):
            mapping[label].append(index)
        return mapping

    def to_set(self):
        return set(label_counts(self._labels)[0].tolist())

    def sort_by_size(self, member_cutoff: Optional[int] = None, max_clusters: Optional[int] = None,
                     bundle=None):
        """Sort labels by cluster size in place, on the GPU (_types.pyx:212-286).

        Cluster 1 becomes the largest non-noise cluster; ties keep the order of
        first occurrence; clusters with fewer than ``member_cutoff`` (default 2)
        members or ranked beyond ``max_clusters`` become noise (0).
        """
        import torch
        from . import engine

        if member_cutoff is None:
            member_cutoff = 2
        if self.n_points == 0:
            return
        old = self._labels
        dev = torch.from_numpy(np.ascontiguousarray(old, dtype=np.int64)).cuda()
        engine.sort_by_size(dev, member_cutoff, max_clusters)
        new = dev.cpu().numpy()
        if self.meta.get("params") or (bundle is not None and getattr(bundle, "_children", None)):
            # old -> new label map for meta["params"] and children (:270-286)
            first = {}
            for o, nl in zip(old.tolist(), new.tolist()):
                if o not in first:
                    first[o] = nl
            params = self.meta.get("params", {})
            self.meta["params"] = {first[k]: v for k, v in params.items() if k in first and first[k] != 0}
            if bundle is not None and getattr(bundle, "_children", None):
                bundle._children = {first[k]: c for k, c in bundle._children.items()
                                    if k in first and first[k] != 0}
        self._labels[:] = new


class ReferenceIndices:
    """Root and parent indices relating a child clustering to its parent
    (_types.pyx:289-302)."""

    def __init__(self, root_indices, parent_indices):
        if root_indices is None or parent_indices is None:
            raise TypeError("root_indices and parent_indices must not be None")
        self._root = root_indices
        self._parent = parent_indices

    @property
    def root(self):
        return np.asarray(self._root)

    @property
    def parent(self):
        return np.asarray(self._parent)


# ---------------------------------------------------------------- input data

class InputData(ABC):
    """Input data interface (_types.pyx:305-335)."""

    @property
    @abstractmethod
    def n_points(self) -> int:
        ...

    @property
    @abstractmethod
    def data(self):
        ...


class InputDataComponents(InputData):
    """Point coordinates (n, d), the equivalent of InputDataExtComponentsMemoryview
    (_types.pyx:1034-1109).

    Host data keeps the reference's float64 contract; a float32 device copy is
    used as-is (its FP64 upcast is exact).  The device tensor is created once
    and cached.
    """

    def __init__(self, data, *, meta=None, device=None):
        import torch

        _meta = {"access_coords": True}
        if meta is not None:
            _meta.update(meta)
        self.meta = _meta
        if isinstance(data, torch.Tensor):
            if data.dim() != 2:
                raise ValueError("coordinates must be 2-D")
            self._host = None
            self._device_tensor = data.contiguous() if data.is_cuda else None
            if not data.is_cuda:
                self._host = np.ascontiguousarray(data.numpy(), dtype=P_AVALUE)
        else:
            arr = np.asarray(data)
            if arr.ndim != 2:
                raise ValueError("coordinates must be 2-D")
            self._host = np.ascontiguousarray(arr, dtype=arr.dtype if arr.dtype in (np.float32, np.float64)
                                              else P_AVALUE)
            self._device_tensor = None
        self._device = device

    @property
    def n_points(self) -> int:
        return int(self._host.shape[0] if self._host is not None else self._device_tensor.shape[0])

    @property
    def n_dim(self) -> int:
        return int(self._host.shape[1] if self._host is not None else self._device_tensor.shape[1])

    @property
    def data(self):
        return self.to_components_array()

    def to_components_array(self):
        if self._host is None:
            self._host = self._device_tensor.cpu().numpy()
        return np.asarray(self._host, dtype=P_AVALUE)

    def device_tensor(self):
        """Contiguous CUDA tensor of the coordinates (float32 or float64)."""
        from . import engine

        if self._device_tensor is None:
            self._device_tensor = engine.to_device_points(self._host, self._device)
        return self._device_tensor

    def get_component(self, point: int, dimension: int) -> float:
        return float(self.to_components_array()[point, dimension])

    def get_subset(self, indices: Sequence) -> "InputDataComponents":
        """Rows ``indices`` of the coordinates, or the sub-matrix ``[indices][:, indices]``
        of a distance matrix (meta["access_dists"]) (_types.pyx:1064-1086)."""
        has_coordinates = self.meta.get("access_coords", False)
        has_distances = self.meta.get("access_dists", False)
        if has_distances and has_coordinates:
            raise RuntimeError(f"{type(self)} can not store coordinates and distances at the same time. "
                               f"Set 'access_coords' or 'access_dists' in {type(self)}.meta accordingly to proceed.")
        if has_distances:
            return type(self)(np.array(self.to_components_array()[indices][:, indices], order="C"))
        if has_coordinates:
            return type(self)(self.to_components_array()[indices])
        raise RuntimeError(f"No information whether coordinates or distances or stored. "
                           f"Set 'access_coords' or 'access_dists' in {type(self)}.meta accordingly to proceed.")

    def by_parts(self):
        """Yield data by parts (meta['edges']) as InputDataExtComponentsMemoryview.by_parts."""
        if self.n_points > 0:
            edges = self.meta.get("edges", None) or [self.n_points]
            data = self.data
            start = 0
            for end in edges:
                yield data[start:(start + end), :]
                start += end


InputDataExtComponentsMemoryview = InputDataComponents


class InputDataNeighbourhoods(InputData):
    """Precomputed neighbourhoods, the equivalent of InputDataExtNeighbourhoodsMemoryview
    (_types.pyx:1153-1217): a padded (n, max_deg) index array plus n_neighbours.

    On the device the lists become a CSR with rows sorted ascending (the
    similarity kernel needs sorted, duplicate-free rows).
    """

    def __init__(self, data, n_neighbours, *, meta=None):
        _meta = {"access_neighbours": True}
        if meta is not None:
            _meta.update(meta)
        self.meta = _meta
        self._data = np.ascontiguousarray(data, dtype=P_AINDEX)
        if self._data.ndim == 1:
            self._data = self._data.reshape(len(n_neighbours), -1)
        self._n_neighbours = np.ascontiguousarray(n_neighbours, dtype=P_AINDEX)
        self._csr = None

    @property
    def n_points(self) -> int:
        return int(self._data.shape[0])

    @property
    def data(self):
        return [s[:self._n_neighbours[i]] for i, s in enumerate(self._data)]

    @property
    def n_neighbours(self):
        return self._n_neighbours

    def get_n_neighbours(self, point: int) -> int:
        return int(self._n_neighbours[point])

    def get_neighbour(self, point: int, member: int) -> int:
        return int(self._data[point, member])

    def get_subset(self, indices: Sequence) -> "InputDataNeighbourhoods":
        """Neighbourhoods of the points ``indices``, each cut to members in
        ``indices``; member indices keep the numbering of the full set, as
        InputDataExtNeighbourhoodsMemoryview.get_subset (_types.pyx:1199-1217).
        The reference filters the padded rows and constructs the subset
        without ``n_neighbours`` (which its constructor requires); here only
        the first ``n_neighbours[i]`` entries are filtered and the new lengths
        are passed."""
        keep = np.zeros(self.n_points, dtype=bool)
        idx = np.asarray(indices, dtype=np.intp)
        keep[idx] = True
        rows = [[int(m) for m in self._data[i, :self._n_neighbours[i]] if keep[m]] for i in idx.tolist()]
        lengths = [len(a) for a in rows]
        width = max(lengths) if lengths else 0
        padded = np.zeros((len(rows), width), dtype=P_AINDEX)
        for i, a in enumerate(rows):
            padded[i, :len(a)] = a
        return type(self)(padded, np.asarray(lengths, dtype=P_AINDEX))

    def device_csr(self, device=None):
        """(row_ptr int64[n+1], col int32[nnz]) on the device, rows sorted ascending."""
        import torch

        if self._csr is None:
            dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
            n = self.n_points
            nn = torch.from_numpy(self._n_neighbours.astype(np.int64)).to(dev)
            pad = torch.from_numpy(self._data.astype(np.int64)).to(dev)
            width = pad.shape[1] if pad.dim() == 2 else 0
            if n and width:
                slot = torch.arange(width, device=dev).unsqueeze(0)
                big = torch.iinfo(torch.int64).max
                pad = torch.where(slot < nn.unsqueeze(1), pad, torch.full_like(pad, big))
                pad, _ = torch.sort(pad, dim=1)
                col = pad[slot.expand(n, width) < nn.unsqueeze(1)].to(torch.int32)
            else:
                col = torch.zeros(0, dtype=torch.int32, device=dev)
            row_ptr = torch.zeros(n + 1, dtype=torch.int64, device=dev)
            if n:
                row_ptr[1:] = torch.cumsum(nn, 0)
            self._csr = (row_ptr, col.contiguous())
        return self._csr


InputDataExtNeighbourhoodsMemoryview = InputDataNeighbourhoods


# ---------------------------------------------------------------- metric

class Metric(ABC):
    """Metric interface (_types.pyx:654-716).

    ``squared_threshold(adjusted_radius)`` is this build's addition: the
    threshold T on the reduced sum s_ij (sequential FP64 sum of rounded squares
    of rounded differences) such that the metric's decision equals s_ij <= T,
    or None when the metric has no such form.
    """

    @abstractmethod
    def adjust_radius(self, radius_cutoff: float) -> float:
        ...

    def squared_threshold(self, adjusted_radius: float):
        return None


class MetricEuclideanReduced(Metric):
    """Squared Euclidean distance (MetricExtEuclideanReduced, _types.pyx:2250-2287).

    The distance itself is evaluated inside the GPU neighbour kernels in the
    reference's order (sequential FP64 sum of separately rounded squares).
    """

    def adjust_radius(self, radius_cutoff: float) -> float:
        r = float(radius_cutoff)
        return r * r

    def squared_threshold(self, adjusted_radius: float):
        return float(adjusted_radius)

    def __str__(self):
        return type(self).__name__


class MetricEuclidean(Metric):
    """Euclidean distance sqrt(s) (MetricExtEuclidean, _types.pyx:2170-2207),
    radius used as given.  fl(sqrt(s)) <= r is decided exactly as s <= T with T
    the largest double whose correctly rounded square root is <= r, so the
    reduced-metric kernels serve it (both precision forms)."""

    def adjust_radius(self, radius_cutoff: float) -> float:
        return float(radius_cutoff)

    def squared_threshold(self, adjusted_radius: float):
        from .engine import sqrt_decision_threshold

        return sqrt_decision_threshold(adjusted_radius)

    def __str__(self):
        return type(self).__name__


class MetricPrecomputed(Metric):
    """Distances read from an (n, n) matrix given as the input data
    (MetricExtPrecomputed, _types.pyx:2105-2127), radius used as given."""

    def adjust_radius(self, radius_cutoff: float) -> float:
        return float(radius_cutoff)

    def __str__(self):
        return type(self).__name__


class MetricEuclideanPeriodicReduced(Metric):
    """Squared Euclidean distance with periodic boundaries
    (MetricExtEuclideanPeriodicReduced, _types.pyx:2290-2348): per dimension
    the difference is reduced modulo its period (``bounds[k] > 0``) and to the
    nearer image.  Radius squared as for the reduced metric."""

    def __init__(self, bounds):
        self._bounds = np.ascontiguousarray(bounds, dtype=np.float64).reshape(-1)

    @property
    def bounds(self):
        return self._bounds

    def adjust_radius(self, radius_cutoff: float) -> float:
        r = float(radius_cutoff)
        return r * r

    def squared_threshold(self, adjusted_radius: float):
        return float(adjusted_radius)

    def __str__(self):
        return f"{type(self).__name__}(bounds={self._bounds.tolist()})"


MetricExtEuclidean = MetricEuclidean
MetricExtPrecomputed = MetricPrecomputed
MetricExtEuclideanPeriodicReduced = MetricEuclideanPeriodicReduced


MetricExtEuclideanReduced = MetricEuclideanReduced


class DistanceGetter(ABC):
    """Distance getter interface (_types.pyx:591-652)."""


class DistanceGetterMetric(DistanceGetter):
    """Forwards to the metric (DistanceGetterExtMetric, _types.pyx:1966-2017)."""

    def __init__(self, metric: Metric = None):
        self._metric = MetricEuclideanReduced() if metric is None else metric

    @classmethod
    def get_builder_kwargs(cls):
        return [("metric", None)]

    def __str__(self):
        return f"{type(self).__name__}(metric={self._metric})"


DistanceGetterExtMetric = DistanceGetterMetric


# ---------------------------------------------------------------- neighbours getters

class NeighboursGetter(ABC):
    """Neighbours getter interface (_types.pyx:478-589)."""

    is_sorted = True
    is_selfcounting = True


class NeighboursGetterGPU(NeighboursGetter):
    """Radius neighbourhoods of all points on the GPU.

    Same sets as NeighboursGetterExtBruteForce._get (_types.pyx:1692-1710):
    N(i) = {j : s_ij <= r2}, i included, rows ascending.  ``precision`` is
    ``"f64"`` (plain FP64 in reference order) or ``"f32_band"`` (FP32 with an
    FP64 recheck inside the guard band; identical sets).
    """

    def __init__(self, distance_getter: DistanceGetter = None, precision: str = "f64"):
        self._distance_getter = DistanceGetterMetric() if distance_getter is None else distance_getter
        self.precision = precision
        _lib.precision_code(precision)
        self.is_sorted = True
        self.is_selfcounting = True

    @classmethod
    def get_builder_kwargs(cls):
        return [("distance_getter", None)]

    def squared_threshold(self, cluster_params: ClusterParameters) -> float:
        """Threshold on the reduced sum for the kernels (r*r for the reduced
        metric); NotImplementedError for metrics without that form."""
        metric = self._distance_getter._metric
        thr = metric.squared_threshold(cluster_params.radius_cutoff)
        if thr is None:
            raise NotImplementedError(f"metric {type(metric).__name__} has no reduced-sum form")
        return thr

    @property
    def precomputed(self) -> bool:
        return isinstance(self._distance_getter._metric, MetricPrecomputed)

    @property
    def periodic_bounds(self):
        metric = self._distance_getter._metric
        return metric.bounds if isinstance(metric, MetricEuclideanPeriodicReduced) else None

    def get_graph(self, input_data: InputDataComponents, cluster_params: ClusterParameters):
        from . import engine

        return engine.build_graph(input_data.device_tensor(), self.squared_threshold(cluster_params),
                                  self.precision, bounds=self.periodic_bounds)

    def get_precomputed_csr(self, input_data: InputDataComponents, cluster_params: ClusterParameters):
        """(row_ptr, col) from a distance matrix given as the input data."""
        from . import engine

        row_ptr, col, asym = engine.precomputed_csr(input_data.device_tensor(), cluster_params.radius_cutoff)
        if asym:
            raise NotImplementedError(f"{asym} point pairs of the distance matrix are neighbours in one "
                                      "direction only; the labelling needs a symmetric neighbour relation")
        return row_ptr, col

    def __str__(self):
        return (f"{type(self).__name__}(dgetter={self._distance_getter}, precision={self.precision}, "
                f"sorted={self.is_sorted}, selfcounting={self.is_selfcounting})")


NeighboursGetterExtBruteForce = NeighboursGetterGPU


class NeighboursGetterLookup(NeighboursGetter):
    """Neighbourhoods read from precomputed lists (NeighboursGetterExtLookup, _types.pyx:1791-1832)."""

    def __init__(self, is_sorted: bool = False, is_selfcounting: bool = True):
        self.is_sorted = is_sorted
        self.is_selfcounting = is_selfcounting

    def __str__(self):
        return f"{type(self).__name__}(sorted={self.is_sorted}, selfcounting={self.is_selfcounting})"


NeighboursGetterExtLookup = NeighboursGetterLookup


# ---------------------------------------------------------------- neighbours containers

class Neighbours(ABC):
    """Neighbours container interface (_types.pyx:438-476)."""


class NeighboursCSR(Neighbours):
    """One row of a neighbour CSR (read-only view).

    ``enough(m)`` is ``n_points > m`` as NeighboursExtVectorCPPUnorderedSet._enough
    (_types.pyx:1561-1564).
    """

    def __init__(self, members=None):
        self._members = np.asarray([] if members is None else members, dtype=P_AINDEX)

    @property
    def n_points(self) -> int:
        return int(self._members.shape[0])

    @property
    def neighbours(self):
        return self._members

    def get_member(self, index: int) -> int:
        return int(self._members[index])

    def contains(self, member: int) -> bool:
        i = np.searchsorted(self._members, member)
        return bool(i < self._members.shape[0] and self._members[i] == member)

    def enough(self, member_cutoff: int) -> bool:
        return self.n_points > member_cutoff


# ---------------------------------------------------------------- similarity checkers

class SimilarityChecker(ABC):
    """Similarity checker interface (_types.pyx:718-780)."""

    sim_mode = _lib.CNN_SIM_AUTO


class SimilarityCheckerGPU(SimilarityChecker):
    """Common-neighbour test of every candidate edge on the GPU.

    Counts |N(i) & N(j)| with early exit at the similarity cutoff, the
    semantics of SimilarityCheckerExtSwitchContains._check / Screensorted._check
    (_types.pyx:2580-2610, 2654-2706).  ``mode``: "merge" (wave-wide merge of
    the sorted lists), "bitset" (AND + popcount of the rows' 64-bit words over
    common column tiles) or "auto" (words when the graph keeps them or the
    lists are dense enough, else the merge).
    """

    def __init__(self, mode: str = "auto"):
        modes = {"auto": _lib.CNN_SIM_AUTO, "merge": _lib.CNN_SIM_MERGE, "bitset": _lib.CNN_SIM_BITSET}
        if mode not in modes:
            raise ValueError(f"unknown similarity mode {mode!r}")
        self.mode = mode
        self.sim_mode = modes[mode]

    def __str__(self):
        return f"{type(self).__name__}(mode={self.mode})"


class PriorityQueueMaxHeap:
    """Max-heap of (a, b, weight) entries ordered by (-weight, a, b)
    (PriorityQueueMaxHeap, _types.pyx:2825-2851).  The MST fitter's native
    forest uses this same order; the class is kept for recipes and callers."""

    def __init__(self):
        self.reset()

    def reset(self) -> None:
        self._queue = []

    def push(self, a, b, weight) -> None:
        heapq.heappush(self._queue, (-weight, (a, b)))

    def pop(self):
        weight, (a, b) = heapq.heappop(self._queue)
        return a, b, -weight

    def is_empty(self) -> bool:
        return not self._queue

    def size(self):
        return len(self._queue)

    def __str__(self):
        return f"{type(self).__name__}"


class QueueUnused:
    """Accepted in the ``fitter.queue`` slot for recipe compatibility.

    The GPU fitter labels by union-find, which gives the BFS result without a
    queue (the FIFO order of QueueExtFIFOQueue does not affect labels).
    """

    def __str__(self):
        return type(self).__name__
