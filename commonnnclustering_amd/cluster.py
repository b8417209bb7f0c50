"""User API: ``Clustering`` (cnnclustering cluster.pyx:102-392).

``Clustering(data, recipe=..., **recipe_kwargs)`` builds the component graph
through the recipe builder; ``fit(radius_cutoff, cnn_cutoff, member_cutoff,
...)`` keeps the reference's contract: cnn_offset subtraction and the
``cnn_cutoff >= 0`` assertion, fresh or frozen label initialisation,
``make_parameters``, timed fitter call, label meta information,
``sort_by_size`` and the ``Record`` appended to the summary.  The hierarchy
methods (``fit_hierarchical``, ``isolate``, ``reel``, the trims and
``to_nx_DiGraph``) work on the bundle tree (cluster.pyx:394-558, 1127-1164).
"""

from __future__ import annotations

import weakref

import numpy as np

from typing import Optional

from ._primitive_types import P_AINDEX
from ._types import InputData, Labels, label_counts
from ._bundle import Bundle
from ._fit import Fitter, FitterGPU, HierarchicalFitter, Predictor
from .report import Record, timed
from . import recipes


class Clustering:
    """Organises a clustering (cluster.pyx:56-147).

    Args:
        data: raw input (2-D sequence, array or CUDA tensor of point
            coordinates; neighbour lists for the neighbourhood recipes), an
            ``InputData`` instance or a ``Bundle``.
        fitter: a ``Fitter`` instance overriding the recipe.
        recipe: a registered recipe name (default ``"coordinates"``) or a
            mapping of component keys to component types.
        **recipe_kwargs: recipe overrides, e.g.
            ``fitter__ngetter__precision=...`` via a tuple spec.
    """

    def __init__(self, data=None, *, fitter=None, hierarchical_fitter=None, predictor=None,
                 bundle_kwargs=None, recipe=None, **recipe_kwargs):
        builder = recipes.Builder(recipe, **recipe_kwargs)
        if bundle_kwargs is None:
            bundle_kwargs = {}
        if data is None:
            self._bundle = None
        elif isinstance(data, Bundle):
            self._bundle = data
        elif isinstance(data, InputData):
            self._bundle = Bundle(data, **bundle_kwargs)
        else:
            self._bundle = Bundle(builder.make_input_data(data), **bundle_kwargs)

        if fitter is not None:
            if not isinstance(fitter, Fitter):
                raise TypeError(f"Can't use object of type {type(fitter).__name__} as fitter. "
                                f"Expected type {Fitter.__name__}.")
            self._fitter = fitter
        else:
            f = builder.make_component("fitter")
            self._fitter = None if f is object else f
        if hierarchical_fitter is not None:
            if not isinstance(hierarchical_fitter, HierarchicalFitter):
                raise TypeError(f"Can't use object of type {type(hierarchical_fitter).__name__} as hierarchical "
                                f"fitter. Expected type {HierarchicalFitter.__name__}.")
            self._hierarchical_fitter = hierarchical_fitter
        else:
            h = builder.make_component("hierarchical_fitter")
            self._hierarchical_fitter = None if h is object else h
        if predictor is not None:
            if not isinstance(predictor, Predictor):
                raise TypeError(f"Can't use object of type {type(predictor).__name__} as predictor. "
                                f"Expected type {Predictor.__name__}.")
            self._predictor = predictor
        else:
            p = builder.make_component("predictor")
            self._predictor = None if p is object else p

    @property
    def root(self):
        return self._bundle

    @property
    def labels(self):
        if self._bundle is None or self._bundle._labels is None:
            return None
        return self._bundle._labels.labels

    @labels.setter
    def labels(self, value):
        if self._bundle is None:
            raise ValueError("Can't set labels because there is no root bundle")
        self._bundle.labels = value

    @property
    def input_data(self):
        if self._bundle is None or self._bundle._input_data is None:
            return None
        return self._bundle._input_data.data

    @property
    def fitter(self):
        return self._fitter

    @fitter.setter
    def fitter(self, value):
        if value is not None and not isinstance(value, Fitter):
            raise TypeError(f"Can't use object of type {type(value).__name__} as fitter. "
                            f"Expected type {Fitter.__name__}.")
        self._fitter = value

    @property
    def children(self):
        """Mapping of child cluster labels to the child bundles of the root."""
        return self._bundle._children

    @property
    def hierarchical_fitter(self):
        return self._hierarchical_fitter

    @property
    def summary(self):
        return self._bundle._summary

    def __getitem__(self, value):
        return self._bundle.get_child(value)

    def __str__(self):
        return (f"{type(self).__name__}(input_data={None if self._bundle is None else self._bundle._input_data}, "
                f"fitter={self._fitter}, hfitter={self._hierarchical_fitter}, predictor={self._predictor})")

    def _fit(self, cluster_params, bundle=None):
        """Low-level fit without parameter preparation (cluster.pyx:230-254)."""
        if bundle is None:
            bundle = self._bundle
        self._fitter.fit(bundle._input_data, bundle._labels, cluster_params)

    def fit(self, radius_cutoff: float, cnn_cutoff: int, member_cutoff: int = None, max_clusters: int = None,
            cnn_offset: int = None, sort_by_size: bool = True, info: bool = True, record: bool = True,
            record_time: bool = True, v: bool = True, purge: bool = False, bundle=None) -> None:
        """Execute the clustering procedure (cluster.pyx:256-392)."""
        if bundle is None:
            bundle = self._bundle
        if cnn_offset is None:
            cnn_offset = 0
        cnn_cutoff -= cnn_offset
        assert cnn_cutoff >= 0

        n = bundle._input_data.n_points
        fresh = (bundle._labels is None) or purge or (not bundle._labels.meta.get("frozen", False))
        if fresh:
            old_label_set = set()
            current_start = 1
        else:
            old_label_set = bundle._labels.to_set()
            current_start = max(old_label_set) + 1

        cluster_params = self._fitter.make_parameters(radius_cutoff, cnn_cutoff, current_start)

        fused = fresh and sort_by_size and isinstance(self._fitter, FitterGPU)
        if fused:
            # Device-resident fit + sort_by_size in one call; the same labels as
            # fit() on fresh Labels followed by Labels.sort_by_size.
            def run():
                return self._fitter.fit_fresh_sorted(bundle._input_data, cluster_params, member_cutoff,
                                                     max_clusters)

            if record_time:
                new_labels, execution_time = timed(run)()
            else:
                new_labels, execution_time = run(), None
            unsorted_params = None
            bundle._labels = Labels(np.ascontiguousarray(new_labels, dtype=P_AINDEX),
                                    consider=np.zeros(n, dtype=np.uint8))
        else:
            if fresh:
                bundle._labels = Labels(np.zeros(n, order="C", dtype=P_AINDEX))
            if record_time:
                _, execution_time = timed(self._fitter.fit)(bundle._input_data, bundle._labels, cluster_params)
            else:
                self._fitter.fit(bundle._input_data, bundle._labels, cluster_params)
                execution_time = None
            unsorted_params = True

        counted = label_counts(bundle._labels.labels) if (info or record) else None
        if info:
            new_label_set = set(counted[0].tolist())
            params = {k: (radius_cutoff, cnn_cutoff) for k in new_label_set - old_label_set if k != 0}
            meta = {"params": params, "reference": weakref.proxy(bundle), "origin": "fit"}
            old_params = bundle._labels.meta.get("params", {})
            old_params.update(meta["params"])
            meta["params"] = old_params
            bundle._labels.meta.update(meta)

        if sort_by_size and unsorted_params:
            bundle._labels.sort_by_size(member_cutoff, max_clusters)

        if record:
            # the Record fields of cluster.pyx:366-390 (a Counter of the labels) from
            # label counts; counted again when sort_by_size ran after the meta step
            uniq, counts = label_counts(bundle._labels.labels) if unsorted_params and sort_by_size else counted
            noise = uniq == 0
            n_noise = int(counts[noise].sum())
            n_largest = int(counts[~noise].max()) if (~noise).any() else 0
            rec = Record(n, radius_cutoff, cnn_cutoff, member_cutoff, max_clusters,
                         int((~noise).sum()), n_largest / n if n else 0.0,
                         n_noise / n if n else 0.0, execution_time)
            if v:
                print(rec)
            bundle._summary.append(rec)

    def fit_hierarchical(self, purge: bool = True, bundle=None, **fitter_kwargs) -> None:
        """Run the hierarchical fitter on a bundle (cluster.pyx:394-411); keyword
        arguments go to its ``fit`` (e.g. radius_cutoff, cnn_cutoff,
        member_cutoff, max_clusters, cnn_offset, sort_by_size, info, v)."""
        if bundle is None:
            bundle = self._bundle
        if self._hierarchical_fitter is None:
            raise AttributeError("this clustering has no hierarchical fitter; pass hierarchical_fitter= "
                                 "or a recipe with one")
        if purge or bundle._children is None:
            bundle._children = {}
        self._hierarchical_fitter.fit(bundle, **fitter_kwargs)

    def isolate(self, purge: bool = True, isolate_input_data: bool = True, bundle=None) -> None:
        """Child clusterings from the cluster labels (cluster.pyx:413-433)."""
        if bundle is None:
            bundle = self._bundle
        bundle.isolate(purge, isolate_input_data)

    def trim_shrinking_leafs(self, bundle=None) -> None:
        """Drop the labels and children of hierarchy nodes below the last split
        on each path (cluster.pyx:435-470): a node is kept if it is new (its
        parent split), splits itself or has a splitting descendant."""
        if bundle is None:
            bundle = self._bundle

        def _trim_shrinking(node, new=True):
            if not node._children:
                splits = will_split = False
            else:
                label_set = node._labels.set
                label_set.discard(0)
                splits = len(label_set) > 1
                will_split = any([_trim_shrinking(child, new=splits) for child in node._children.values()])
            keep = new or will_split or splits
            if not keep:
                node._labels = None
                node._children = {}
            return keep

        _trim_shrinking(bundle)

    def trim_trivial_leafs(self, bundle=None) -> None:
        """Reset labels and children of nodes whose labels are all noise
        (cluster.pyx:472-500)."""
        if bundle is None:
            bundle = self._bundle

        def _trim_trivial(node):
            if node._labels is None:
                return
            if node._labels.set == {0}:
                node._labels = None
                node._children = {}
                return
            if not node._children:
                return
            for child in node._children.values():
                _trim_trivial(child)

        _trim_trivial(bundle)

    def reel(self, depth: Optional[int] = None, bundle=None) -> None:
        """Fold the labels of lower hierarchy levels into their parents
        (cluster.pyx:502-558): for each labelled child (recursively, up to
        ``depth`` levels), its points get the child's labels shifted past the
        parent's current maximum label (noise stays 0), and the child's
        ``params`` move to the parent under the shifted labels."""
        if bundle is None:
            bundle = self._bundle

        def _reel(parent, depth):
            if not parent._children:
                return
            if depth is not None:
                depth -= 1
            parent._labels.meta["origin"] = "reel"
            parent_labels = parent._labels.labels
            for label, child in parent._children.items():
                if child._labels is None:
                    continue
                if depth is None or depth > 0:
                    _reel(child, depth)
                n_clusters = max(parent_labels)
                child_labels = child._labels.labels
                parent_labels[child.parent_indices] = np.where(child_labels == 0, 0, child_labels + n_clusters)
                try:
                    _ = parent._labels.meta["params"].pop(label)
                except KeyError:
                    pass
                params = child._labels.meta.get("params", {})
                for old_label, p in params.items():
                    parent._labels.meta["params"][old_label + n_clusters] = p

        if depth is not None:
            assert depth > 0
        _reel(bundle, depth)

    def to_nx_DiGraph(self, ignore=None, bundle=None):
        """The cluster hierarchy as a networkx DiGraph (cluster.pyx:1127-1164):
        nodes are dotted label paths from "1" (the root), each with the bundle
        as attribute ``object``; children in ascending label order."""
        import networkx as nx

        if bundle is None:
            bundle = self._bundle
        ignore = set() if ignore is None else set(ignore)

        def add_children(node_label, node, graph):
            for child_label, child in sorted(node._children.items()):
                if child_label in ignore:
                    continue
                padded = ".".join([node_label, str(child_label)])
                graph.add_node(padded, object=child)
                graph.add_edge(node_label, padded)
                if child._children:
                    add_children(padded, child, graph)

        graph = nx.DiGraph()
        graph.add_node("1", object=bundle)
        add_children("1", bundle, graph)
        return graph

    def predict(self, other, radius_cutoff: float, cnn_cutoff: int, clusters=None, cnn_offset: int = None,
                info: bool = True, record: bool = True, record_time: bool = True, v: bool = True,
                purge: bool = False, bundle=None) -> None:
        """Predict cluster labels for the points of ``other`` (a ``Bundle`` or a
        ``Clustering``) from this clustering's labels (cluster.pyx:560-653).

        Args:
            other: the data to label.
            radius_cutoff: neighbour search radius.
            cnn_cutoff: similarity criterion (as in ``fit``).
            clusters: labels that may be given (default: every label of this
                clustering except 0).
            cnn_offset: subtracted from ``cnn_cutoff`` (compatibility).
            purge: re-initialise the labels of ``other`` (they are also reset
                when missing or not frozen).
        """
        if bundle is None:
            bundle = self._bundle
        if isinstance(other, Clustering):
            other = other._bundle
        if self._predictor is None:
            raise AttributeError("this clustering has no predictor; pass predictor= or use a recipe with one")
        if cnn_offset is None:
            cnn_offset = 0
        cnn_cutoff -= cnn_offset
        assert cnn_cutoff >= 0

        if (other._labels is None) or purge or (not other._labels.meta.get("frozen", False)):
            other._labels = Labels(np.zeros(other._input_data.n_points, order="C", dtype=P_AINDEX))

        cluster_params = self._predictor.make_parameters(radius_cutoff, cnn_cutoff, 0)
        if clusters is None:
            clusters = bundle._labels.to_set() - {0}
        other._labels.consider_set = clusters

        args = (bundle._input_data, other._input_data, bundle._labels, other._labels, cluster_params)
        if record_time:
            _, execution_time = timed(self._predictor.predict)(*args)
        else:
            self._predictor.predict(*args)
            execution_time = None

        if info:
            params = {k: (radius_cutoff, cnn_cutoff) for k in clusters if k != 0}
            meta = {"params": params, "reference": weakref.proxy(bundle), "origin": "predict"}
            old_params = other._labels.meta.get("params", {})
            old_params.update(meta["params"])
            meta["params"] = old_params
            other._labels.meta.update(meta)
        other._labels.meta["frozen"] = True

    def to_dtrajs(self, bundle=None):
        """Labels split into the parts the input was given in (``meta["edges"]``),
        as Clustering.to_dtrajs (cluster.pyx:1166-1188): one array per part, a
        trailing empty part dropped; the whole array when there are no edges."""
        if bundle is None:
            bundle = self._bundle
        labels_array = bundle.labels
        if labels_array is None:
            return []
        edges = None
        if bundle._input_data is not None:
            edges = bundle._input_data.meta.get("edges")
        if edges is None:
            return [labels_array]
        dtrajs = np.split(labels_array, np.cumsum(edges))
        if len(dtrajs[-1]) == 0:
            dtrajs = dtrajs[:-1]
        return dtrajs
