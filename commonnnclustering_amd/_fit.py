"""Fitters: the Fitter interface and the GPU fitter.

``Fitter`` mirrors the reference ABC (_fit.pyx:43-85), including
``make_parameters`` (r -> metric-adjusted radius; self-counting getters turn
c into c+2 similarity / c+1 member cutoffs, _fit.pyx:129-161).  ``FitterGPU``
replaces ``FitterExtBFS`` (_fit.pyx:356-504): same ``fit(input_data, labels,
cluster_params)`` contract (labels and consider mutated in place), computed as
GPU neighbour search + similarity test + union-find.
"""

from __future__ import annotations

import weakref
from abc import ABC, abstractmethod
from collections.abc import Iterable

import numpy as np

from . import _lib
from ._primitive_types import P_AINDEX
from ._types import (
    ClusterParameters,
    InputDataComponents,
    InputDataNeighbourhoods,
    Labels,
    NeighboursGetterGPU,
    NeighboursGetterLookup,
    SimilarityCheckerGPU,
)


class Fitter(ABC):
    """Fitter interface (_fit.pyx:43-85)."""

    @abstractmethod
    def fit(self, input_data, labels: Labels, cluster_params: ClusterParameters):
        """Generic clustering"""

    def make_parameters(self, radius_cutoff: float, cnn_cutoff: int, current_start: int) -> ClusterParameters:
        try:
            used_metric = self._neighbours_getter._distance_getter._metric
        except AttributeError:
            pass
        else:
            radius_cutoff = used_metric.adjust_radius(radius_cutoff)

        n_member_cutoff = cnn_cutoff
        try:
            is_selfcounting = self._neighbours_getter.is_selfcounting
        except AttributeError:
            pass
        else:
            if is_selfcounting:
                n_member_cutoff += 1
                cnn_cutoff += 2

        return ClusterParameters(radius_cutoff, cnn_cutoff, cnn_cutoff, n_member_cutoff, current_start)


def _to_host(t, owner):
    """Copy a device int64 vector to a new host array through a pinned
    staging buffer cached on ``owner`` (a pageable copy runs at a fraction of
    the link rate)."""
    import torch

    n = t.numel()
    buf = getattr(owner, "_pinned", None)
    if buf is None or buf.numel() < n:
        buf = torch.empty(max(n, 1), dtype=t.dtype, pin_memory=True)
        owner._pinned = buf
    view = buf[:n]
    view.copy_(t)
    return view.numpy().copy()


class FitterGPU(Fitter):
    """Common-nearest-neighbours fit on the GPU.

    Produces the labels of FitterExtBFS._fit (_fit.pyx:408-491): clusters are
    the connected components of (core points, passing edges), numbered from
    ``current_start`` by ascending smallest point index; points with
    ``consider == 0`` are skipped (but still counted inside neighbourhoods);
    every consider flag is 0 afterwards.

    Args:
        neighbours_getter: ``NeighboursGetterGPU`` (coordinates) or
            ``NeighboursGetterLookup`` (precomputed neighbourhoods).
        similarity_checker: ``SimilarityCheckerGPU``.
        neighbours, neighbour_neighbours, queue: accepted for recipe
            compatibility with FitterExtBFS.get_builder_kwargs (_fit.pyx:398-406);
            unused by the GPU procedure.
    """

    def __init__(self, neighbours_getter=None, neighbours=None, neighbour_neighbours=None,
                 similarity_checker=None, queue=None):
        self._neighbours_getter = NeighboursGetterGPU() if neighbours_getter is None else neighbours_getter
        self._neighbours = neighbours
        self._neighbour_neighbours = neighbour_neighbours
        self._similarity_checker = SimilarityCheckerGPU() if similarity_checker is None else similarity_checker
        self._queue = queue
        self.last_graph_stats = None
        self.last_label_stats = None

    @classmethod
    def get_builder_kwargs(cls):
        return [
            ("neighbours_getter", None),
            ("neighbours", None),
            ("neighbour_neighbours", "neighbours"),
            ("similarity_checker", None),
            ("queue", None),
        ]

    def __str__(self):
        return (f"{type(self).__name__}(ngetter={self._neighbours_getter}, "
                f"checker={self._similarity_checker})")

    # -- device helpers ---------------------------------------------------
    def _sim_mode(self):
        return getattr(self._similarity_checker, "sim_mode", _lib.CNN_SIM_AUTO)

    def fit(self, input_data, labels: Labels, cluster_params: ClusterParameters):
        import torch
        from . import engine

        n = input_data.n_points
        if labels.n_points != n:
            raise ValueError("labels and input data differ in length")
        if n == 0:
            return
        dev = torch.device("cuda", torch.cuda.current_device())
        lab = torch.from_numpy(np.ascontiguousarray(labels._labels, dtype=np.int64)).to(dev)
        con = torch.from_numpy(np.ascontiguousarray(labels._consider, dtype=np.uint8)).to(dev)
        c_sim = cluster_params.similarity_cutoff
        m_core = cluster_params.n_member_cutoff
        if isinstance(input_data, InputDataNeighbourhoods) or isinstance(self._neighbours_getter,
                                                                        NeighboursGetterLookup):
            row_ptr, col = input_data.device_csr(dev)
            _, st = engine.labels_from_csr(row_ptr, col, c_sim, m_core, consider=con,
                                           current_start=cluster_params.current_start, labels=lab,
                                           sim_mode=self._sim_mode())
            self.last_graph_stats = None
        elif self._neighbours_getter.precomputed:
            row_ptr, col = self._neighbours_getter.get_precomputed_csr(input_data, cluster_params)
            _, st = engine.labels_from_csr(row_ptr, col, c_sim, m_core, consider=con,
                                           current_start=cluster_params.current_start, labels=lab,
                                           sim_mode=self._sim_mode())
            self.last_graph_stats = None
        else:
            graph = self._neighbours_getter.get_graph(input_data, cluster_params)
            try:
                _, st = graph.labels(c_sim, m_core, consider=con, current_start=cluster_params.current_start,
                                     labels=lab, sim_mode=self._sim_mode())
                self.last_graph_stats = graph.stats
            finally:
                graph.release()
        self.last_label_stats = st
        labels._labels[:] = lab.cpu().numpy()
        labels._consider[:] = con.cpu().numpy()

    def fit_fresh_sorted(self, input_data, cluster_params: ClusterParameters, member_cutoff, max_clusters):
        """Fused fresh fit + sort_by_size on the device; returns int64 labels (host).

        Equivalent to fit() on zero labels with all points considered followed
        by Labels.sort_by_size(member_cutoff, max_clusters).
        """
        from . import engine

        if isinstance(input_data, InputDataNeighbourhoods) or isinstance(self._neighbours_getter,
                                                                        NeighboursGetterLookup):
            import torch

            row_ptr, col = input_data.device_csr()
            lab = torch.zeros(input_data.n_points, dtype=torch.int64, device=row_ptr.device)
            _, st = engine.labels_from_csr(row_ptr, col, cluster_params.similarity_cutoff,
                                           cluster_params.n_member_cutoff, current_start=1, labels=lab,
                                           sim_mode=self._sim_mode())
            engine.sort_by_size(lab, member_cutoff, max_clusters)
            self.last_graph_stats, self.last_label_stats = None, st
            return _to_host(lab, self)
        if self._neighbours_getter.precomputed:
            import torch

            row_ptr, col = self._neighbours_getter.get_precomputed_csr(input_data, cluster_params)
            lab = torch.zeros(input_data.n_points, dtype=torch.int64, device=row_ptr.device)
            _, st = engine.labels_from_csr(row_ptr, col, cluster_params.similarity_cutoff,
                                           cluster_params.n_member_cutoff, current_start=1, labels=lab,
                                           sim_mode=self._sim_mode())
            engine.sort_by_size(lab, member_cutoff, max_clusters)
            self.last_graph_stats, self.last_label_stats = None, st
            return _to_host(lab, self)
        if self._neighbours_getter.periodic_bounds is not None:
            graph = self._neighbours_getter.get_graph(input_data, cluster_params)
            try:
                lab, ls = graph.fit(cluster_params.similarity_cutoff, cluster_params.n_member_cutoff, sort=True,
                                    member_cutoff=member_cutoff, max_clusters=max_clusters)
                self.last_graph_stats, self.last_label_stats = graph.stats, ls
            finally:
                graph.release()
            return _to_host(lab, self)
        lab, _, gs, ls = engine.fit(input_data.device_tensor(), self._neighbours_getter.squared_threshold(cluster_params),
                                    cluster_params.similarity_cutoff, cluster_params.n_member_cutoff,
                                    precision=self._neighbours_getter.precision, sort=True,
                                    member_cutoff=member_cutoff, max_clusters=max_clusters)
        self.last_graph_stats, self.last_label_stats = gs, ls
        return _to_host(lab, self)


    def fit_steps(self, input_data, params_list, sort_by_size: bool = True, member_cutoff=None,
                  max_clusters=None):
        """Fresh labels (host int64 arrays) for a list of ClusterParameters.

        Each entry equals fit() on zero labels with every point considered
        and current_start 1, followed by Labels.sort_by_size(member_cutoff,
        max_clusters) when ``sort_by_size``.  Steps of one radius share one
        graph; smaller radii are cut from the graph of the largest by exact
        FP64 re-thresholding (not for the periodic metric, whose graphs are
        built per radius).
        """
        import torch
        from . import engine

        results = [None] * len(params_list)
        if not params_list:
            return results
        if isinstance(input_data, InputDataNeighbourhoods) or isinstance(self._neighbours_getter,
                                                                        NeighboursGetterLookup) \
                or self._neighbours_getter.precomputed:
            csr = {}
            for k, p in enumerate(params_list):
                key = None if isinstance(input_data, InputDataNeighbourhoods) or isinstance(
                    self._neighbours_getter, NeighboursGetterLookup) else p.radius_cutoff
                if key not in csr:
                    csr[key] = (input_data.device_csr() if key is None
                                else self._neighbours_getter.get_precomputed_csr(input_data, p))
                row_ptr, col = csr[key]
                lab = torch.zeros(input_data.n_points, dtype=torch.int64, device=row_ptr.device)
                engine.labels_from_csr(row_ptr, col, p.similarity_cutoff, p.n_member_cutoff, current_start=1,
                                       labels=lab, sim_mode=self._sim_mode())
                if sort_by_size:
                    engine.sort_by_size(lab, member_cutoff, max_clusters)
                results[k] = lab.cpu().numpy()
            return results
        ng = self._neighbours_getter
        by_thr = {}
        for k, p in enumerate(params_list):
            by_thr.setdefault(ng.squared_threshold(p), []).append(k)
        periodic = ng.periodic_bounds is not None
        base = None
        try:
            for thr in sorted(by_thr, reverse=True):
                steps = by_thr[thr]
                if base is None or periodic:
                    g = ng.get_graph(input_data, params_list[steps[0]])
                    if base is None:
                        base = g
                else:
                    g = base.rethreshold(thr)
                try:
                    for k in steps:
                        p = params_list[k]
                        lab, _ = g.fit(p.similarity_cutoff, p.n_member_cutoff, sort=sort_by_size,
                                       member_cutoff=member_cutoff, max_clusters=max_clusters)
                        results[k] = lab.cpu().numpy()
                finally:
                    if g is not base:
                        g.release()
        finally:
            if base is not None:
                base.release()
        return results

    def neighbour_csr(self, input_data, cluster_params: ClusterParameters):
        """Neighbour lists at the radius of ``cluster_params``: (row_ptr int64[n+1],
        col int32[nnz]) on the device, caller order, rows ascending."""
        return neighbour_csr(self._neighbours_getter, input_data, cluster_params)


def neighbour_csr(neighbours_getter, input_data, cluster_params: ClusterParameters):
    """Device CSR of the neighbourhoods NeighboursGetter.get gives for every
    point (caller order, rows ascending, self included for coordinates)."""
    if isinstance(input_data, InputDataNeighbourhoods) or isinstance(neighbours_getter, NeighboursGetterLookup):
        return input_data.device_csr()
    if neighbours_getter.precomputed:
        return neighbours_getter.get_precomputed_csr(input_data, cluster_params)
    graph = neighbours_getter.get_graph(input_data, cluster_params)
    try:
        return graph.to_csr()
    finally:
        graph.release()


FitterExtBFS = FitterGPU


# ---------------------------------------------------------------- hierarchical fitters

class HierarchicalFitter(ABC):
    """Hierarchical fitter interface (_fit.pyx:88-126).

    ``make_parameters`` looks up ``self._neighbours_getter`` as the fitter's
    does; a hierarchical fitter without one (HierarchicalFitterRepeat) gets
    the radius and cnn_cutoff unchanged, as in the reference.
    """

    @abstractmethod
    def fit(self, bundle, **kwargs):
        """Generic hierarchical clustering"""

    make_parameters = Fitter.make_parameters


def _as_list(v, cast):
    if not isinstance(v, Iterable) or isinstance(v, (str, bytes)):
        v = [v]
    return [cast(x) for x in v]


class HierarchicalFitterRepeat(HierarchicalFitter):
    """Hierarchy from repeated fits (HierarchicalFitterRepeat, _fit.pyx:1014-1187).

    Step k fits all points at (radius_cutoff[k], cnn_cutoff[k]) (a scalar is
    repeated to the other list's length) on fresh labels, then sorts them by
    size.  The clusters of step k become children of the step-(k-1) cluster
    holding their points: each terminal bundle of the previous step gets as
    labels its points' step-k labels and is isolated (without input data);
    the new children are the next terminal bundles.  The step fits run on
    the GPU with one graph per radius.

    ``make_parameters`` passes the radius and cnn_cutoff unchanged (this
    fitter has no neighbours getter; the reference's does the same), so the
    step fits compare reduced distances with ``radius_cutoff`` itself and use
    ``cnn_cutoff`` as similarity and member cutoff.

    The reference visits the parents of a step in std::unordered_map order;
    here they are visited in ascending label.  The order matters only when
    one cluster has points under several parents (possible when a later step
    merges clusters, e.g. a growing radius): the parent visited last holds
    the terminal bundle of that cluster, and the next step's isolate of it
    fails with an IndexError, as the reference's does.  Steps that only split
    clusters (radius not growing, cnn_cutoff not falling) never do this.
    """

    def __init__(self, fitter=None):
        self._fitter = FitterGPU() if fitter is None else fitter

    def __str__(self):
        return f"{type(self).__name__}(fitter={self._fitter})"

    @classmethod
    def get_builder_kwargs(cls):
        return [("fitter", None)]

    def fit(self, bundle, **kwargs):
        self._fit(bundle, kwargs["radius_cutoff"], kwargs["cnn_cutoff"], kwargs.get("member_cutoff"),
                  kwargs.get("max_clusters"), kwargs.get("cnn_offset"), kwargs.get("sort_by_size", True),
                  kwargs.get("info", True), kwargs.get("v", True))

    def _step_labels(self, input_data, params_list, sort_by_size, member_cutoff, max_clusters):
        if hasattr(self._fitter, "fit_steps"):
            return self._fitter.fit_steps(input_data, params_list, sort_by_size, member_cutoff, max_clusters)
        out = []
        for p in params_list:
            lab = Labels(np.zeros(input_data.n_points, order="C", dtype=P_AINDEX))
            self._fitter.fit(input_data, lab, p)
            if sort_by_size:
                lab.sort_by_size(member_cutoff, max_clusters)
            out.append(lab.labels.copy())
        return out

    def _fit(self, bundle, radius_cutoff, cnn_cutoff, member_cutoff=None, max_clusters=None, cnn_offset=None,
             sort_by_size=True, info=True, v=True):
        from ._bundle import label_groups

        radius_cutoff = _as_list(radius_cutoff, float)
        cnn_cutoff = _as_list(cnn_cutoff, int)
        if len(radius_cutoff) == 1:
            radius_cutoff *= len(cnn_cutoff)
        if len(cnn_cutoff) == 1:
            cnn_cutoff *= len(radius_cutoff)
        n_steps = len(radius_cutoff)
        assert n_steps == len(cnn_cutoff)
        if cnn_offset is None:
            cnn_offset = 0
        n = bundle._input_data.n_points
        params_list = [self.make_parameters(radius_cutoff[k], cnn_cutoff[k] - cnn_offset, 1)
                       for k in range(n_steps)]
        step_labels = self._step_labels(bundle._input_data, params_list, sort_by_size, member_cutoff,
                                        max_clusters)

        previous = np.ones(n, order="C", dtype=P_AINDEX)
        terminal = {1: bundle}
        for step in range(n_steps):
            if v:
                print(f"Running step {step:<5} (r = {radius_cutoff[step]}, c = {cnn_cutoff[step]})")
            current = np.ascontiguousarray(step_labels[step], dtype=P_AINDEX)
            new_terminal = {}
            for p_label, points in label_groups(previous, order_by="label"):
                if p_label == 0:
                    continue
                parent = terminal[p_label]
                parent._labels = Labels.from_sequence(current[points])
                if info:
                    params = {k: (radius_cutoff[step], cnn_cutoff[step]) for k in parent._labels.to_set() if k != 0}
                    parent._labels.meta.update({"params": params, "reference": weakref.proxy(parent),
                                                "origin": "fit"})
                parent.isolate(isolate_input_data=False)
                for c_label, child in parent._children.items():
                    if c_label == 0:
                        continue
                    new_terminal[c_label] = child
            terminal = new_terminal
            previous = current


class HierarchicalFitterMSTPrim(HierarchicalFitter):
    """MST cluster hierarchy (HierarchicalFitterMSTPrim, _fit.pyx:733-1011).

    The neighbour graph at ``radius_cutoff`` is weighted by common-neighbour
    counts and its maximum spanning forest is grown by Prim's procedure with
    the reference's queue order (weights and forest in the native library:
    cnn_similarity_mst).  The tree edges are then merged in order of
    decreasing weight into a hierarchy of bundles whose ``_graph`` holds the
    merged edges and whose ``_lambda`` is the weight of the merging edge;
    children are pruned and folded by ``check_children`` with
    ``member_cutoff`` (default 10).  Labels of the root are all 1 with every
    consider flag 0, as the reference leaves them.

    Args:
        neighbours_getter: ``NeighboursGetterGPU`` (coordinates, precomputed
            distances) or ``NeighboursGetterLookup``.
        neighbours, neighbour_neighbours, similarity_checker, priority_queue,
        priority_queue_tree: accepted for recipe compatibility
            (_fit.pyx:782-790); the queues' order is built into the native
            forest.
    """

    def __init__(self, neighbours_getter=None, neighbours=None, neighbour_neighbours=None,
                 similarity_checker=None, priority_queue=None, priority_queue_tree=None):
        self._neighbours_getter = NeighboursGetterGPU() if neighbours_getter is None else neighbours_getter
        self._neighbours = neighbours
        self._neighbour_neighbours = neighbour_neighbours
        self._similarity_checker = SimilarityCheckerGPU() if similarity_checker is None else similarity_checker
        self._priority_queue = priority_queue
        self._priority_queue_tree = priority_queue_tree
        self.last_tree = None

    def __str__(self):
        return (f"{type(self).__name__}(ngetter={self._neighbours_getter}, na={self._neighbours}, "
                f"nb={self._neighbour_neighbours}, checker={self._similarity_checker}, "
                f"prioq={self._priority_queue}, prioq (tree)={self._priority_queue_tree})")

    @classmethod
    def get_builder_kwargs(cls):
        return [
            ("neighbours_getter", None),
            ("neighbours", None),
            ("neighbour_neighbours", "neighbours"),
            ("similarity_checker", None),
            ("priority_queue", None),
            ("priority_queue_tree", "priority_queue"),
        ]

    def fit(self, bundle, **kwargs):
        self._fit(bundle, kwargs["radius_cutoff"], kwargs.get("member_cutoff"), kwargs.get("max_clusters"),
                  kwargs.get("cnn_offset"), kwargs.get("sort_by_size", True), kwargs.get("info", True),
                  kwargs.get("v", True))

    def spanning_forest(self, input_data, radius_cutoff: float):
        """Tree edges (a, b, weight) as host int64 arrays sorted by (-weight, a, b)."""
        from . import engine

        params = self.make_parameters(radius_cutoff, 0, 1)
        row_ptr, col = neighbour_csr(self._neighbours_getter, input_data, params)
        return engine.similarity_mst(row_ptr, col)

    def _fit(self, bundle, radius_cutoff, member_cutoff=None, max_clusters=None, cnn_offset=None,
             sort_by_size=True, info=True, v=True):
        try:
            import networkx  # noqa: F401
        except ModuleNotFoundError:  # pragma: no cover
            raise ModuleNotFoundError("No module named 'networkx'") from None
        from ._bundle import build_mst_hierarchy

        n = bundle._input_data.n_points
        labels = Labels(np.ones(n, order="C", dtype=P_AINDEX))
        bundle._labels = labels
        if member_cutoff is None:
            member_cutoff = 10
        a, b, w = self.spanning_forest(bundle._input_data, radius_cutoff)
        labels._consider[:] = 0
        self.last_tree = (a, b, w)
        build_mst_hierarchy(bundle, n, a.tolist(), b.tolist(), w.tolist(), member_cutoff)


class Predictor(ABC):
    """Predictor interface (_fit.pyx:164-216): ``predict`` labels the points of
    a second data set from a labelled one; ``make_parameters`` as the fitter's."""

    @abstractmethod
    def predict(self, input_data, predictand_input_data, labels: Labels, predictand_labels: Labels,
                cluster_params: ClusterParameters):
        """Generic cluster label prediction"""

    make_parameters = Fitter.make_parameters


class PredictorGPU(Predictor):
    """First-match prediction on the GPU, replacing PredictorFirstmatch.predict
    (_fit.pyx:1194-1291) with brute-force getters: for every considered
    predictand point with more than n_member_cutoff reference points within
    the radius, the reference neighbours are visited in ascending index and the
    first whose label is in ``predictand_labels.consider_set``, which has more
    than n_member_cutoff neighbours itself and which shares at least
    similarity_cutoff neighbours with the point gives it its label (its
    consider flag becomes 0).

    Args:
        neighbours_getter: ``NeighboursGetterGPU``; its ``precision`` selects the
            distance form (identical neighbour sets either way).
        neighbours_getter_other, neighbours, neighbour_neighbours,
        similarity_checker: accepted for recipe compatibility with
            PredictorFirstmatch.get_builder_kwargs (_fit.pyx:1222-1230).
    """

    def __init__(self, neighbours_getter=None, neighbours_getter_other=None, neighbours=None,
                 neighbour_neighbours=None, similarity_checker=None):
        self._neighbours_getter = NeighboursGetterGPU() if neighbours_getter is None else neighbours_getter
        self._neighbours_getter_other = (self._neighbours_getter if neighbours_getter_other is None
                                         else neighbours_getter_other)
        self._neighbours = neighbours
        self._neighbour_neighbours = neighbour_neighbours
        self._similarity_checker = SimilarityCheckerGPU() if similarity_checker is None else similarity_checker
        self.last_tests = None

    @classmethod
    def get_builder_kwargs(cls):
        return [
            ("neighbours_getter", None),
            ("neighbours_getter_other", "neighbours_getter"),
            ("neighbours", None),
            ("neighbour_neighbours", "neighbours"),
            ("similarity_checker", None),
        ]

    def __str__(self):
        return (f"{type(self).__name__}(ngetter={self._neighbours_getter}, "
                f"checker={self._similarity_checker})")

    def predict(self, input_data, predictand_input_data, labels: Labels, predictand_labels: Labels,
                cluster_params: ClusterParameters):
        import torch
        from . import engine

        if not isinstance(input_data, InputDataComponents) or not isinstance(predictand_input_data,
                                                                             InputDataComponents):
            raise NotImplementedError("GPU prediction needs point coordinates for both data sets")
        if predictand_labels.n_points != predictand_input_data.n_points:
            raise ValueError("predictand labels and input data differ in length")
        if getattr(self._neighbours_getter, "periodic_bounds", None) is not None:
            raise NotImplementedError("prediction under the periodic metric is not offered: the reference's "
                                      "_calc_distance_other for it sums (a_k - b_k)^2 without the periodic "
                                      "reduction (_types.pyx:2332-2343)")
        X = input_data.device_tensor()
        Y = predictand_input_data.device_tensor()
        dev = X.device
        lab_x = torch.from_numpy(np.ascontiguousarray(labels._labels, dtype=np.int64)).to(dev)
        lab_y = torch.from_numpy(np.ascontiguousarray(predictand_labels._labels, dtype=np.int64)).to(dev)
        con_y = torch.from_numpy(np.ascontiguousarray(predictand_labels._consider, dtype=np.uint8)).to(dev)
        precision = getattr(self._neighbours_getter, "precision", "f64")
        _, _, tests = engine.predict(X, lab_x, Y.to(dev), self._neighbours_getter.squared_threshold(cluster_params),
                                     cluster_params.similarity_cutoff, cluster_params.n_member_cutoff,
                                     predictand_labels.consider_set, labels_y=lab_y, consider_y=con_y,
                                     precision=precision)
        self.last_tests = tests
        predictand_labels._labels[:] = lab_y.cpu().numpy()
        predictand_labels._consider[:] = con_y.cpu().numpy()


PredictorFirstmatch = PredictorGPU
