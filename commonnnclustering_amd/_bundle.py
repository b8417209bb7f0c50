"""Bundles: one clustering's input data, labels, reference indices and child
clusterings, and the tree operations on them (cnnclustering _bundle.pyx).

A ``Bundle`` is a node of the cluster hierarchy.  A child made by
``isolate`` holds the points of one cluster of its parent: ``parent_indices``
index the parent's points, ``root_indices`` the points of the root bundle.
"""

from __future__ import annotations

import weakref
from collections import deque
from collections.abc import Iterable, MutableMapping, MutableSequence

import numpy as np

from ._types import InputData, Labels, ReferenceIndices
from .report import Summary


def label_groups(labels: np.ndarray, order_by: str = "first"):
    """[(label, ascending point indices)] in order of each label's first
    occurrence (``order_by="first"``: the items of Labels.to_mapping,
    _types.pyx:187-198) or of ascending label (``"label"``), grouped with one
    stable sort instead of a loop over the points (a radix sort when the
    labels span fewer than 2^16 values)."""
    labels = np.asarray(labels)
    n = labels.shape[0]
    if n == 0:
        return []
    lo, hi = int(labels.min()), int(labels.max())
    key = (labels - lo).astype(np.uint16) if hi - lo < 65536 else labels
    order = np.argsort(key, kind="stable")
    sk = key[order]
    starts = np.flatnonzero(np.r_[True, sk[1:] != sk[:-1]])
    ends = np.r_[starts[1:], n]
    groups = [(int(labels[order[a]]), order[a:b]) for a, b in zip(starts.tolist(), ends.tolist())]
    if order_by == "label":
        return groups
    return [groups[k] for k in np.argsort(order[starts], kind="stable")]


class Bundle:
    """Node of a cluster hierarchy (_bundle.pyx:20-272).

    Args:
        input_data: an ``InputData`` instance (or None).
        graph: the hierarchy graph of an MST-built node (a networkx graph).
        labels: ``Labels`` or a label sequence.
        reference_indices: ``ReferenceIndices`` relating this node's points to
            its parent's and the root's.
        children: mapping of cluster label to child ``Bundle``.
        alias: name (default ``"root"``).
        parent: the parent bundle (kept as a weak proxy); sets
            ``hierarchy_level`` to the parent's plus one.
    """

    def __init__(self, input_data=None, graph=None, labels=None, reference_indices=None, children=None,
                 alias=None, parent=None, meta=None, summary=None, hierarchy_level=0):
        self.input_data = input_data
        self._graph_build = None
        self._graph_n = None
        self._graph = graph
        self.labels = labels
        self._reference_indices = reference_indices
        self.children = children
        self.alias = "root" if alias is None else str(alias)
        if parent is not None:
            self._parent = weakref.proxy(parent)
            self.hierarchy_level = parent.hierarchy_level + 1
        else:
            self._parent = None
            self.hierarchy_level = hierarchy_level
        self.meta = {} if meta is None else meta
        self.summary = summary
        self._lambda = -np.inf

    # -- hierarchy graph: a value, or built on first access
    @property
    def _graph(self):
        if self._graph_build is not None:
            build, self._graph_build = self._graph_build, None
            self._graph_value = build()
        return self._graph_value

    @_graph.setter
    def _graph(self, value):
        self._graph_build = None
        self._graph_n = None
        self._graph_value = value

    def _set_graph_builder(self, build, n_nodes=None):
        """Defer the hierarchy graph: ``build()`` makes it on first access;
        ``n_nodes`` answers its node count until then."""
        self._graph_value = None
        self._graph_build = build
        self._graph_n = n_nodes

    def _graph_len(self) -> int:
        """len(self._graph), without building a deferred graph whose node count is known."""
        if self._graph_build is not None and self._graph_n is not None:
            return self._graph_n
        g = self._graph
        return 0 if g is None else len(g)

    # -- properties (_bundle.pyx:54-140)
    @property
    def labels(self):
        return None if self._labels is None else self._labels.labels

    @labels.setter
    def labels(self, value):
        if value is not None and not isinstance(value, Labels):
            value = Labels(value)
        self._labels = value

    @property
    def root_indices(self):
        return None if self._reference_indices is None else self._reference_indices.root

    @property
    def parent_indices(self):
        return None if self._reference_indices is None else self._reference_indices.parent

    @property
    def input_data(self):
        return None if self._input_data is None else self._input_data.data

    @input_data.setter
    def input_data(self, value):
        if value is not None and not isinstance(value, InputData):
            raise TypeError(f"Can't use object of type {type(value).__name__} as input data. "
                            f"Expected type {InputData.__name__}.")
        self._input_data = value

    @property
    def children(self):
        """Mapping of child cluster labels to child bundles."""
        return self._children

    @children.setter
    def children(self, value):
        if value is None:
            value = {}
        elif type(value) is not dict and not isinstance(value, MutableMapping):
            raise TypeError("Expected a mutable mapping")
        self._children = value

    @property
    def hierarchy_level(self):
        """Level in the hierarchy (0 for the root)."""
        return self._hierarchy_level

    @hierarchy_level.setter
    def hierarchy_level(self, value):
        self._hierarchy_level = int(value)

    @property
    def summary(self):
        return self._summary

    @summary.setter
    def summary(self, value):
        if value is None:
            value = Summary()
        elif not isinstance(value, MutableSequence):
            raise TypeError("Expected a mutable sequence")
        self._summary = value

    def info(self):
        """Alias, level, kind of data access, point count and child count (_bundle.pyx:148-169)."""
        access = []
        if self._input_data is not None:
            for kind, check in (("coordinates", "access_coords"), ("distances", "access_distances"),
                                ("neighbours", "access_neighbours")):
                if self._input_data.meta.get(check, False):
                    access.append(kind)
        if not access:
            access = ["unknown"]
        n_points = self._input_data.n_points if self._input_data is not None else None
        return "\n".join([
            f"alias: {self.alias!r}",
            f"hierarchy_level: {self._hierarchy_level}",
            f"access: {', '.join(access)}",
            f"points: {n_points}",
            f"children: {len(self.children)}",
        ])

    # -- child access (_bundle.pyx:171-260)
    @staticmethod
    def _label_path(label):
        """An int, a list of ints or a dotted string ("1.1.2") as an int or a list."""
        if isinstance(label, str):
            label = label.split(".")
        if isinstance(label, Iterable):
            label = [int(x) for x in label]
            if len(label) == 1:
                label = label[0]
        return label

    def get_child(self, label):
        """The child with this label, or the last one along a label path."""
        label = self._label_path(label)
        if isinstance(label, (int, np.integer)):
            try:
                return self._children[int(label)]
            except KeyError:
                raise KeyError(f"Clustering {self.alias!r} has no child with label {label}") from None
        next_label, *rest = label
        try:
            child = self._children[next_label]
        except KeyError:
            raise KeyError(f"Clustering {self.alias!r} has no child with label {next_label}") from None
        return child.get_child(rest)

    def add_child(self, label, bundle=None):
        """Add (or silently replace) a child; ``bundle=None`` makes a new empty one.

        A plain int label inserts into this bundle's children; the reference
        (_bundle.pyx:249-254) inserts and then fails unpacking the int, while
        its own test fixtures rely on the insertion, which is what this does.
        """
        if bundle is None:
            bundle = type(self)(parent=self)
        if not isinstance(bundle, Bundle):
            raise TypeError("child must be a Bundle")
        label = self._label_path(label)
        if isinstance(label, (int, np.integer)):
            self._children[int(label)] = bundle
            return
        *rest, last = label
        self.get_child(rest)._children[last] = bundle

    def __getitem__(self, key):
        return self.get_child(key)

    def __setitem__(self, key, value):
        self.add_child(key, value)

    def isolate(self, purge: bool = True, isolate_input_data: bool = True):
        """One child per cluster label (see ``isolate``)."""
        isolate(self, purge, isolate_input_data)


def _child_edges(edges, parent_indices: np.ndarray):
    """Points of a child per part of the parent's input (meta["edges"]), as
    the walk of _bundle.pyx:329-350: one count per part of the parent."""
    bounds = np.cumsum(np.asarray(edges, dtype=np.int64))
    part = np.searchsorted(bounds, parent_indices, side="right")
    return np.bincount(part, minlength=len(edges))[:len(edges)].tolist()


def isolate(bundle: Bundle, purge: bool = True, isolate_input_data: bool = True) -> None:
    """Child bundles from cluster labels (_bundle.pyx:275-352).

    For every label except 0 (in order of first occurrence) a child with the
    cluster's point indices: ``parent_indices`` into this bundle, and
    ``root_indices`` through this bundle's own root indices.  With
    ``isolate_input_data`` the child gets the subset of the input data, and
    when the input has parts (``meta["edges"]``) the child's points per part.
    """
    if purge or bundle._children is None:
        bundle._children = {}
    for label, indices in label_groups(bundle._labels.labels):
        if label == 0:
            continue
        parent_indices = np.asarray(indices, dtype=np.intp)
        if bundle._reference_indices is None:
            root_indices = parent_indices
        else:
            root_indices = bundle._reference_indices.root[parent_indices]
        child = Bundle(parent=bundle)
        child._reference_indices = ReferenceIndices(root_indices, parent_indices)
        parent_alias = bundle.alias if bundle.alias is not None else ""
        # the reference appends to the child's default alias ("root")
        child.alias += f"{parent_alias} - {label}"
        bundle._children[label] = child
        if not isolate_input_data:
            continue
        child._input_data = bundle._input_data.get_subset(parent_indices)
        edges = bundle._input_data.meta.get("edges", None)
        if edges is None:
            continue
        child._input_data.meta["edges"] = _child_edges(edges, parent_indices) if edges else []


def check_children(bundle: Bundle, member_cutoff: int, needs_folding: bool = False) -> None:
    """Prune and fold the children of an MST hierarchy node (_bundle.pyx:355-395).

    With ``needs_folding`` children merged at this node's own lambda are
    replaced by their first descendants at another lambda.  Children with
    fewer than ``member_cutoff`` graph nodes are dropped and the rest are
    renumbered from 1; a single remaining child is replaced by its children
    (and this node takes its lambda).
    """
    if needs_folding:
        leafs = []
        queue = deque()
        for child in bundle._children.values():
            (queue if child._lambda == bundle._lambda else leafs).append(child)
        while queue:
            candidate = queue.popleft()
            for child in candidate._children.values():
                (queue if child._lambda == bundle._lambda else leafs).append(child)
        bundle._children = {k: child for k, child in enumerate(leafs, 1)}

    bundle._children = {k: v for k, v in enumerate(bundle._children.values(), 1) if v._graph_len() >= member_cutoff}

    if len(bundle._children) == 1:
        child = bundle._children.popitem()[1]
        for label, grandchild in enumerate(child._children.values(), 1):
            bundle._children[label] = grandchild
            bundle._lambda = child._lambda

    for child in bundle._children.values():
        child._parent = weakref.proxy(bundle)


def bfs_leafs(root: Bundle):
    """The bundles of a hierarchy, breadth first from ``root`` (_bundle.pyx:398-409)."""
    yield root
    q = deque([root])
    while q:
        bundle = q.popleft()
        for child in bundle._children.values():
            yield child
            q.append(child)


def build_mst_hierarchy(bundle: Bundle, n_points: int, tree_a, tree_b, tree_w, member_cutoff: int) -> None:
    """Bundle hierarchy from spanning-forest edges sorted by (-weight, a, b),
    as the second half of HierarchicalFitterMSTPrim._fit (_fit.pyx:918-1011).

    The edges are merged in order into growing point sets (union-find): each
    merge makes a top bundle holding the merging edge plus everything the (up
    to two) top bundles it joins hold, which become its children (1 and 2);
    its ``_lambda`` is the edge weight.  Before the first edge of a lower
    weight, every live top bundle not yet checked is pruned by
    ``check_children`` (folding children of equal lambda when the previous
    edges tied).  A forest leaves the root with lambda -inf, an empty graph
    and the live top bundles of at least ``member_cutoff`` nodes as children; a single tree
    gives the root the graph and lambda of the last top bundle (its children
    are left as they were, as in the reference).

    The reference copies every merged subtree's graph at each merge (time
    and memory of the order of points x depth).  Here a top bundle's graph
    (networkx, edge attribute ``weight``) is built on first access from the
    merge records, and pruning uses node counts kept with the union-find;
    the graphs hold the same nodes and weighted edges as the reference's
    (their insertion order can differ).
    """
    import gc

    # one bundle per edge: the cyclic collector's passes over them would dominate
    enabled = gc.isenabled()
    gc.disable()
    try:
        _build_mst_hierarchy(bundle, n_points, tree_a, tree_b, tree_w, member_cutoff)
    finally:
        if enabled:
            gc.enable()


def _build_mst_hierarchy(bundle, n_points, tree_a, tree_b, tree_w, member_cutoff):
    import networkx as nx

    parent = list(range(n_points))
    size = [1] * n_points
    set_label = list(range(n_points))  # per set root: n_points + top index, or the point itself
    rec_a, rec_b, rec_w, rec_l, rec_r = [], [], [], [], []
    top_bundles = {}
    pending = []  # tops created since the last check, in creation order

    def find(p):
        root = p
        while parent[root] != root:
            root = parent[root]
        while parent[p] != root:  # path compression: the set roots stay the same
            parent[p], p = root, parent[p]
        return root

    def subtree_edges(tops):
        out, stack = [], list(reversed(tops))
        while stack:
            k = stack.pop()
            out.append((rec_a[k], rec_b[k], {"weight": rec_w[k]}))
            if rec_r[k] >= 0:
                stack.append(rec_r[k])
            if rec_l[k] >= 0:
                stack.append(rec_l[k])
        return out

    def graph_of(tops):
        return lambda: nx.Graph(subtree_edges(tops))

    def check_pending(folding):
        for index in pending:
            top = top_bundles.get(index)
            if top is not None:
                check_children(top, member_cutoff, folding)
        pending.clear()

    current_weight = np.inf
    needs_folding = False
    i = -1
    for i, (a, b, weight) in enumerate(zip(tree_a, tree_b, tree_w)):
        ra, rb = find(a), find(b)
        la, lb = set_label[ra], set_label[rb]
        if weight < current_weight:
            check_pending(needs_folding)
            needs_folding = False
        else:
            needs_folding = True

        top = Bundle()
        top._lambda = weight
        top_bundles[i] = top
        pending.append(i)
        rec_a.append(a), rec_b.append(b), rec_w.append(weight)
        label = 1
        left = right = -1
        if la >= n_points:
            left = la - n_points
            top._children[label] = top_bundles.pop(left)
            label += 1
        if lb >= n_points:
            right = lb - n_points
            top._children[label] = top_bundles.pop(right)
        rec_l.append(left), rec_r.append(right)
        parent[ra] = rb
        size[rb] += size[ra]
        top._set_graph_builder(graph_of([i]), size[rb])
        set_label[rb] = i + n_points
        current_weight = weight

    check_pending(needs_folding)

    if i < n_points - 2:
        bundle._lambda = -np.inf
        count = 1
        for top in top_bundles.values():
            if top._graph_len() >= member_cutoff:
                bundle._children[count] = top
                count += 1
        # the reference adds the root's own, empty, edge set to itself
        # (_fit.pyx:1001): a forest's root keeps an empty graph
        bundle._set_graph_builder(nx.Graph, 0)
        for child in bundle._children.values():
            child._parent = weakref.proxy(bundle)
    else:
        top = top_bundles[n_points - 2]
        bundle._set_graph_builder(lambda: top._graph, top._graph_len())
        bundle._lambda = top._lambda
