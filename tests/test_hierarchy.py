"""Cluster hierarchy: bundle tree operations (host code) against the
reference's own test expectations, and the oracle's hierarchy restatements
against mathematics that fixes them (no GPU)."""

import numpy as np
import pytest

from commonnnclustering_amd import Clustering
from commonnnclustering_amd._bundle import Bundle, build_mst_hierarchy, check_children, label_groups
from commonnnclustering_amd._types import InputDataComponents, InputDataNeighbourhoods, Labels, ReferenceIndices
from oracle import cnn_oracle as O
from kat_data import kats

K = kats()


def _report(clustering, labels):
    """The text test_isolate (test/unit/test_clustering.py:122-160) writes."""
    out = ""
    for label in sorted(set(labels)):
        child = clustering._bundle._children[label]
        points = child._input_data
        out += (f"Child {label}\n{'=' * 80}\nData:\n{points.data}\nEdges:\n{points.meta.get('edges', 'None')}\n"
                f"Root:\n{child.root_indices}\nParent:\n{child.parent_indices}\n\n")
    return out


@pytest.mark.parametrize("k", range(4))
def test_isolate_matches_reference_reports(k):
    case = K["isolate"]["cases"][k]
    data = InputDataComponents(np.array(case["data"], dtype=np.float64), meta=case["meta"])
    cl = Clustering(data=data, bundle_kwargs={"labels": Labels(np.array(case["labels"], dtype=np.int64))})
    if case["root_indices"] is not None:
        cl._bundle._reference_indices = ReferenceIndices(np.array(case["root_indices"]),
                                                         np.array(case["parent_indices"]))
    cl.isolate()
    assert len(cl._bundle._children) == len(set(case["labels"]))
    for label in set(case["labels"]):
        assert cl._bundle._children[label]._input_data.n_points == case["labels"].count(label)
    assert _report(cl, case["labels"]) == case["report"]


def _hierarchical_a():
    """The fixture hierarchy of test/conftest.py (make_hierarchical_clustering_a)."""
    r = K["reel"]
    bundle = Bundle(labels=Labels(np.array(r["root_labels"], dtype=np.int64)))
    cl = Clustering(bundle)
    for path in r["empty_children"]:
        (bundle.get_child(path[:-1]) if len(path) > 1 else bundle).add_child(path[-1])
    for node in r["nodes"]:
        child = bundle.get_child(node["path"])
        child._labels = Labels(np.array(node["labels"], dtype=np.int64))
        child._reference_indices = ReferenceIndices(np.array(node["root"]), np.array(node["parent"]))
    return cl


@pytest.mark.parametrize("case", K["reel"]["cases"], ids=lambda c: f"depth{c['depth']}")
def test_reel_matches_reference(case):
    cl = _hierarchical_a()
    cl.reel(depth=case["depth"])
    np.testing.assert_array_equal(cl._bundle.labels, case["expected"])
    assert cl._bundle._labels.meta["origin"] == "reel"


def test_child_paths():
    cl = _hierarchical_a()
    b = cl._bundle
    assert b.get_child("1.2") is b.get_child([1, 2]) is cl["1.2"] is b[1][2]
    assert b.get_child(1).hierarchy_level == 1 and b.get_child([1, 2]).hierarchy_level == 2
    new = Bundle()
    b["1.2.7"] = new
    assert b.get_child([1, 2, 7]) is new
    with pytest.raises(KeyError):
        b.get_child(9)
    with pytest.raises(KeyError):
        b.get_child("1.9")
    assert "children: 3" in b.info()


def test_isolate_edges_and_reference_chain():
    # parts of 2, 0 and 3 points; a grandchild's root indices go through its parent's
    data = InputDataComponents(np.arange(10.0).reshape(5, 2), meta={"edges": [2, 0, 3]})
    b = Bundle(data, labels=Labels(np.array([1, 0, 2, 1, 1])))
    b.isolate()
    assert list(b.children) == [1, 2]
    assert b[1]._input_data.meta["edges"] == [1, 0, 2]
    assert b[2]._input_data.meta["edges"] == [0, 0, 1]
    np.testing.assert_array_equal(b[1].root_indices, [0, 3, 4])
    b[1].labels = [0, 5, 5]
    b[1].isolate()
    np.testing.assert_array_equal(b[1][5].parent_indices, [1, 2])
    np.testing.assert_array_equal(b[1][5].root_indices, [3, 4])
    np.testing.assert_array_equal(b[1][5].input_data, np.arange(10.0).reshape(5, 2)[[3, 4]])
    assert b[1][5]._input_data.meta["edges"] == [0, 0, 2]
    # the reference appends "<parent alias> - <label>" to the child's default alias "root"
    assert b[1].alias == "rootroot - 1" and b[1][5].alias == "rootrootroot - 1 - 5"


def test_label_groups_first_occurrence_order():
    lab = np.array([3, 0, 3, 1, 0, 7, 1])
    groups = label_groups(lab)
    ref = Labels(lab).to_mapping()
    assert [g[0] for g in groups] == list(ref.keys())
    for label, idx in groups:
        assert idx.tolist() == ref[label]


def test_neighbourhood_subset_keeps_member_numbering():
    data = InputDataNeighbourhoods(np.array([[0, 1, 2], [1, 0, 0], [2, 0, 3], [3, 2, 0]]),
                                   np.array([3, 2, 3, 2]))
    sub = data.get_subset([0, 2, 3])
    assert [list(x) for x in sub.data] == [[0, 2], [2, 0, 3], [3, 2]]


def test_trim_trivial_and_shrinking():
    b = Bundle(labels=Labels(np.array([1, 1, 2, 2, 0])))
    cl = Clustering(b)
    b.isolate(isolate_input_data=False)
    b[1].labels = [0, 0]          # all noise: trivial
    b[2].labels = [3, 3]          # one cluster only: shrinks
    b[2].isolate(isolate_input_data=False)
    b[2][3].labels = [4, 5]       # splits below a non-splitting node
    b[2][3].isolate(isolate_input_data=False)
    cl.trim_trivial_leafs()
    assert b[1]._labels is None and b[1].children == {}
    assert b[2]._labels is not None
    cl.trim_shrinking_leafs()
    # node 2 does not split, but its descendant does: kept, as is the splitting node
    assert b[2]._labels is not None and b[2][3]._labels is not None
    # the new leaves of the split are kept (their parent splits)
    assert set(b[2][3].children) == {4, 5}


def test_trim_shrinking_drops_chain_without_split():
    b = Bundle(labels=Labels(np.array([1, 1, 2, 2])))
    cl = Clustering(b)
    b.isolate(isolate_input_data=False)
    b[1].labels = [6, 6]
    b[1].isolate(isolate_input_data=False)
    b[1][6].labels = [7, 7]
    cl.trim_shrinking_leafs()
    # root splits: child 1 is new and kept; below it nothing splits, so child 6 is dropped
    assert b[1]._labels is not None
    assert b[1][6]._labels is None and b[1][6].children == {}


def test_check_children_prunes_and_lifts():
    import networkx as nx

    def node(edges, lam, children=None):
        x = Bundle(graph=nx.Graph([(a, c, {"weight": lam}) for a, c in edges]))
        x._lambda = lam
        if children:
            x._children = dict(enumerate(children, 1))
        return x

    small = node([(0, 1)], 5)
    big = node([(2, 3), (3, 4), (4, 5)], 5)
    parent = node([(0, 1), (2, 3), (3, 4), (4, 5), (1, 2)], 2, [small, big])
    check_children(parent, member_cutoff=3)
    # the small child is dropped; the single survivor is replaced by its (no) children
    assert parent.children == {}
    assert parent._lambda == 2
    g1, g2 = node([(6, 7), (7, 8)], 9), node([(9, 10), (10, 11)], 9)
    big2 = node([(2, 3), (3, 4), (4, 5)], 5, [g1, g2])
    parent2 = node([(0, 1)], 2, [small, big2])
    check_children(parent2, member_cutoff=3)
    assert parent2.children == {1: g1, 2: g2} and parent2._lambda == 5


# ---------------------------------------------------------------- oracle pins

def _points(n, seed, d=2):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-3, 3, size=(4, d))
    return np.ascontiguousarray(centres[rng.integers(0, 4, n)] + rng.normal(scale=0.5, size=(n, d)))


def _walk(node, path=()):
    yield path, node
    for k, child in node["children"].items():
        yield from _walk(child, path + (k,))


@pytest.mark.parametrize("steps", [([1.0], [2, 4, 8]), ([1.2, 0.8, 0.5], [3]), ([0.9, 0.9], [1, 6])])
def test_repeat_oracle_nodes_are_step_fits(steps):
    """Each node's labels at depth k are the step-k fit restricted to the
    node's points (the definition of the Repeat hierarchy), computed here by an
    independent fit through the components restatement."""
    X = _points(120, 3)
    radii, cuts = steps
    root = O.hierarchical_repeat(X, radii, cuts, member_cutoff=3)
    n_steps = max(len(radii), len(cuts))
    radii = radii * n_steps if len(radii) == 1 else radii
    cuts = cuts * n_steps if len(cuts) == 1 else cuts
    D = O.reduced_distances(X, X)
    step_labels = []
    for r, c in zip(radii, cuts):
        lists = [np.nonzero(row <= r)[0] for row in D]
        row_ptr = np.r_[0, np.cumsum([len(x) for x in lists])]
        lab = np.zeros(len(X), dtype=np.int64)
        O.fit_components(row_ptr, np.concatenate(lists), {"similarity_cutoff": c, "n_member_cutoff": c,
                                                          "current_start": 1}, lab, np.ones(len(X), np.uint8))
        O.sort_by_size_np(lab, 3)
        step_labels.append(lab)
    seen = 0
    for path, node in _walk(root):
        if node["labels"] is None:
            continue
        depth = len(path)
        pts = np.arange(len(X)) if node["root"] is None else node["root"]
        np.testing.assert_array_equal(node["labels"], step_labels[depth][pts])
        if depth:
            parent_labels = step_labels[depth - 1][pts]
            assert set(parent_labels.tolist()) == {path[-1]}
        seen += 1
    assert seen > 1


def _scipy_max_forest_weight(n, neigh, weights):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree

    top = max((max(w) for w in weights if w), default=0) + 1
    rows, cols, vals = [], [], []
    for a in range(n):
        for t, b in enumerate(neigh[a]):
            if b > a:
                rows.append(a), cols.append(b), vals.append(top - weights[a][t])
    m = minimum_spanning_tree(coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr())
    m = m.tocoo()
    return sum(top - v for v in m.data), len(m.data), set(zip(np.minimum(m.row, m.col).tolist(),
                                                               np.maximum(m.row, m.col).tolist()))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_prim_oracle_is_a_maximum_spanning_forest(seed):
    X = _points(150, seed)
    neigh = O.neighbourhoods(X, 0.6)
    w = O.similarity_weights(neigh)
    for a in range(len(neigh)):     # weights are symmetric and count both endpoints
        for t, b in enumerate(neigh[a]):
            assert w[a][t] == len(set(neigh[a]) & set(neigh[b])) >= 2 or a == b
    tree = O.mst_prim(neigh, w)
    total, m, _ = _scipy_max_forest_weight(len(X), neigh, w)
    assert len(tree) == m and sum(e[2] for e in tree) == total
    keys = [(-e[2], e[0], e[1]) for e in tree]
    assert keys == sorted(keys)


def test_prim_oracle_distinct_weights_edge_set():
    rng = np.random.default_rng(7)
    n = 60
    neigh = [[] for _ in range(n)]
    wt = {}
    perm = rng.permutation(10_000)
    k = 0
    for a in range(n):
        for b in range(a + 1, n):
            if rng.random() < 0.15:
                neigh[a].append(b), neigh[b].append(a)
                wt[(a, b)] = int(perm[k]) + 1
                k += 1
    neigh = [sorted(x) for x in neigh]
    weights = [[wt[(min(a, b), max(a, b))] for b in neigh[a]] for a in range(n)]
    tree = O.mst_prim(neigh, weights)
    _, _, edges = _scipy_max_forest_weight(n, neigh, weights)
    assert {(min(a, b), max(a, b)) for a, b, _ in tree} == edges


def _canon_oracle(node):
    return {"lambda": node["lambda"], "nodes": sorted(node["nodes"]),
            "edges": sorted((u, v, w) for (u, v), w in node["edges"].items()),
            "children": {k: _canon_oracle(c) for k, c in node["children"].items()}}


def canon_bundle(b):
    g = b._graph
    nodes = sorted(g.nodes) if g is not None else []
    edges = sorted((min(u, v), max(u, v), d["weight"]) for u, v, d in g.edges(data=True)) if g is not None else []
    return {"lambda": b._lambda, "nodes": nodes, "edges": edges,
            "children": {k: canon_bundle(c) for k, c in b._children.items()}}


def _random_tree_edges(n, seed, distinct):
    rng = np.random.default_rng(seed)
    edges = []
    for b in range(1, n):
        a = int(rng.integers(0, b))
        w = int(rng.integers(1, 10_000)) if distinct else int(rng.integers(1, 5))
        edges.append((a, b, w))
    if distinct:
        ws = rng.permutation(np.arange(1, n))
        edges = [(a, b, int(w)) for (a, b, _), w in zip(edges, ws)]
    edges.sort(key=lambda e: (-e[2], e[0], e[1]))
    return edges


@pytest.mark.parametrize("distinct", [True, False])
@pytest.mark.parametrize("member_cutoff", [2, 4])
def test_mst_hierarchy_builder_matches_oracle(distinct, member_cutoff):
    for seed in range(5):
        n = 40
        edges = _random_tree_edges(n, seed, distinct)
        if seed % 2:  # a forest: drop a few edges
            edges = [e for k, e in enumerate(edges) if k % 9 != 3]
        root = Bundle()
        build_mst_hierarchy(root, n, [e[0] for e in edges], [e[1] for e in edges], [e[2] for e in edges],
                            member_cutoff)
        expect = O.mst_hierarchy(n, edges, member_cutoff)
        expect.pop("top", None)
        got = canon_bundle(root)
        if len(edges) == n - 1:
            got["children"] = expect["children"] = {}  # the reference leaves a single tree's root childless
        assert got == _canon_oracle(expect)


def _leaf(lam, nodes, edges):
    return {"lambda": lam, "nodes": nodes, "edges": edges, "children": {}}


# Hand-worked runs of the merge loop of _fit.pyx:945-1011 with check_children
# (_bundle.pyx:354-398); point n-1 stays isolated, so the root is a forest root:
# lambda -inf, an empty graph (the no-op at _fit.pyx:1001), its children the
# live top bundles of at least member_cutoff nodes.
MST_KATS = [
    # Weights 5, 5, 4, 3, 3: the last top bundle (lambda 3) has two children, the
    # (0..3) bundle of the same lambda and the lambda-4 pair (4, 5).  Folding (the
    # previous edge tied) replaces the same-lambda child by its children: the
    # lambda-4 pair first (a direct child), then the two lambda-5 pairs.
    (7, [(0, 1, 5), (2, 3, 5), (4, 5, 4), (1, 2, 3), (3, 4, 3)], 2,
     {1: {"lambda": 3, "nodes": [0, 1, 2, 3, 4, 5],
          "edges": [(0, 1, 5), (1, 2, 3), (2, 3, 5), (3, 4, 3), (4, 5, 4)],
          "children": {1: _leaf(4, [4, 5], [(4, 5, 4)]), 2: _leaf(5, [0, 1], [(0, 1, 5)]),
                       3: _leaf(5, [2, 3], [(2, 3, 5)])}}}),
    # Distinct weights 6, 5, 4, 3: no folding; the last top bundle keeps one child
    # (the lambda-4 bundle), whose two children are lifted and whose lambda (4)
    # the parent takes over.
    (6, [(0, 1, 6), (2, 3, 5), (1, 2, 4), (3, 4, 3)], 2,
     {1: {"lambda": 4, "nodes": [0, 1, 2, 3, 4], "edges": [(0, 1, 6), (1, 2, 4), (2, 3, 5), (3, 4, 3)],
          "children": {1: _leaf(6, [0, 1], [(0, 1, 6)]), 2: _leaf(5, [2, 3], [(2, 3, 5)])}}}),
    # Weights 5, 4, 3 in a chain, member_cutoff 3: the lambda-4 bundle loses its
    # 2-node child; the last bundle keeps that bundle alone, which has no children
    # to lift, so its lambda stays 3 and it ends childless.
    (5, [(0, 1, 5), (1, 2, 4), (2, 3, 3)], 3,
     {1: _leaf(3, [0, 1, 2, 3], [(0, 1, 5), (1, 2, 4), (2, 3, 3)])}),
]


@pytest.mark.parametrize("n,edges,member_cutoff,children", MST_KATS)
def test_mst_hierarchy_hand_worked(n, edges, member_cutoff, children):
    expect = {"lambda": -np.inf, "nodes": [], "edges": [], "children": children}
    assert _canon_oracle(O.mst_hierarchy(n, edges, member_cutoff)) == expect
    root = Bundle()
    build_mst_hierarchy(root, n, [e[0] for e in edges], [e[1] for e in edges], [e[2] for e in edges],
                        member_cutoff)
    assert canon_bundle(root) == expect



def test_mst_hierarchy_oracle_nodes_are_single_linkage_clusters():
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    n = 50
    for seed in range(4):
        edges = _random_tree_edges(n, seed, True)
        if seed % 2:
            edges = edges[:-7]
        root = O.mst_hierarchy(n, edges, member_cutoff=2)
        # the components of the tree edges of weight >= t, for every edge weight t
        clusters = set()
        for t in sorted({w for _, _, w in edges}):
            keep = [(a, b) for a, b, w in edges if w >= t]
            m = coo_matrix((np.ones(len(keep)), ([a for a, _ in keep], [b for _, b in keep])), shape=(n, n))
            _, comp = connected_components(m, directed=False)
            for c in np.unique(comp):
                members = np.flatnonzero(comp == c)
                if len(members) > 1:
                    clusters.add(tuple(members.tolist()))
        checked = 0
        for _, node in _walk(root.get("top", root)):
            if node is root:
                continue
            assert tuple(sorted(node["nodes"])) in clusters
            for child in node["children"].values():
                assert child["nodes"] < node["nodes"]
            checked += 1
        assert checked > 3
