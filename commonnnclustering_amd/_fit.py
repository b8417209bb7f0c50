"""Fitters: the Fitter interface and the GPU fitter.

``Fitter`` mirrors the reference ABC (_fit.pyx:43-85), including
``make_parameters`` (r -> metric-adjusted radius; self-counting getters turn
c into c+2 similarity / c+1 member cutoffs, _fit.pyx:129-161).  ``FitterGPU``
replaces ``FitterExtBFS`` (_fit.pyx:356-504): same ``fit(input_data, labels,
cluster_params)`` contract (labels and consider mutated in place), computed as
GPU neighbour search + similarity test + union-find.
"""

from __future__ import annotations

from abc import ABC, abstractmethod

import numpy as np

from . import _lib
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
            return lab.cpu().numpy()
        if self._neighbours_getter.precomputed:
            import torch

            row_ptr, col = self._neighbours_getter.get_precomputed_csr(input_data, cluster_params)
            lab = torch.zeros(input_data.n_points, dtype=torch.int64, device=row_ptr.device)
            _, st = engine.labels_from_csr(row_ptr, col, cluster_params.similarity_cutoff,
                                           cluster_params.n_member_cutoff, current_start=1, labels=lab,
                                           sim_mode=self._sim_mode())
            engine.sort_by_size(lab, member_cutoff, max_clusters)
            self.last_graph_stats, self.last_label_stats = None, st
            return lab.cpu().numpy()
        if self._neighbours_getter.periodic_bounds is not None:
            graph = self._neighbours_getter.get_graph(input_data, cluster_params)
            try:
                lab, ls = graph.fit(cluster_params.similarity_cutoff, cluster_params.n_member_cutoff, sort=True,
                                    member_cutoff=member_cutoff, max_clusters=max_clusters)
                self.last_graph_stats, self.last_label_stats = graph.stats, ls
            finally:
                graph.release()
            return lab.cpu().numpy()
        lab, _, gs, ls = engine.fit(input_data.device_tensor(), self._neighbours_getter.squared_threshold(cluster_params),
                                    cluster_params.similarity_cutoff, cluster_params.n_member_cutoff,
                                    precision=self._neighbours_getter.precision, sort=True,
                                    member_cutoff=member_cutoff, max_clusters=max_clusters)
        self.last_graph_stats, self.last_label_stats = gs, ls
        return lab.cpu().numpy()


FitterExtBFS = FitterGPU


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
