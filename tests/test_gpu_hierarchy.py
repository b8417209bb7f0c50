"""Cluster hierarchies on the GPU against the oracle: similarity weights and
the Prim forest (native), the MST bundle hierarchy and the Repeat hierarchy
(HierarchicalFitterRepeat) node by node."""

import numpy as np
import pytest
import torch

from oracle import cnn_oracle as O
from commonnnclustering_amd import Clustering, engine
from commonnnclustering_amd._fit import FitterGPU, HierarchicalFitterMSTPrim, HierarchicalFitterRepeat
from commonnnclustering_amd._lib import CNNError
from commonnnclustering_amd._types import NeighboursGetterGPU
from test_hierarchy import canon_bundle, _canon_oracle

pytestmark = pytest.mark.gpu


def _blobs(n, d, seed, k=4, spread=0.6):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-4, 4, size=(k, d))
    X = centres[rng.integers(0, k, size=n)] + rng.normal(0, spread, size=(n, d))
    return np.ascontiguousarray(X)


def _csr(X, r, precision="f64"):
    dtype = torch.float32 if precision == "f32_band" else torch.float64
    Xd = torch.from_numpy(X).to(dtype).cuda()
    g = engine.build_graph(Xd, r * r, precision)
    try:
        return g.to_csr()
    finally:
        g.release()


@pytest.mark.parametrize("d,n,r,seed", [(2, 700, 0.5, 0), (3, 900, 0.8, 1), (16, 600, 3.2, 2), (64, 400, 6.5, 3)])
def test_edge_similarity_matches_oracle(cuda, d, n, r, seed):
    X = _blobs(n, d, seed)
    neigh = O.neighbourhoods(X, r)
    row_ptr, col = _csr(X, r)
    assert col.cpu().numpy().tolist() == np.concatenate(neigh).tolist()
    w = engine.edge_similarity(row_ptr, col).cpu().numpy()
    want = np.concatenate([np.asarray(x, dtype=np.int64) for x in O.similarity_weights(neigh)])
    np.testing.assert_array_equal(w, want)


def test_edge_similarity_rejects_unsorted_rows(cuda):
    row_ptr = torch.tensor([0, 2, 3], dtype=torch.int64, device="cuda")
    col = torch.tensor([1, 0, 1], dtype=torch.int32, device="cuda")
    with pytest.raises(CNNError, match="ascending"):
        engine.edge_similarity(row_ptr, col)


@pytest.mark.parametrize("d,n,r,seed", [(2, 500, 0.45, 4), (3, 600, 0.7, 5), (16, 400, 3.0, 6)])
def test_similarity_mst_matches_oracle_prim(cuda, d, n, r, seed):
    X = _blobs(n, d, seed)
    neigh = O.neighbourhoods(X, r)
    want = O.mst_prim(neigh)
    a, b, w = engine.similarity_mst(*_csr(X, r))
    assert list(zip(a.tolist(), b.tolist(), w.tolist())) == want


@pytest.mark.parametrize("member_cutoff", [2, 5, 10])
@pytest.mark.parametrize("d,n,r,seed", [(2, 400, 0.35, 7), (3, 350, 0.6, 8)])
def test_mst_hierarchy_matches_oracle(cuda, member_cutoff, d, n, r, seed):
    X = _blobs(n, d, seed)
    tree = O.mst_prim(O.neighbourhoods(X, r))
    expect = O.mst_hierarchy(n, tree, member_cutoff)
    expect.pop("top", None)
    cl = Clustering(X, hfitter="mst")
    assert isinstance(cl.hierarchical_fitter, HierarchicalFitterMSTPrim)
    cl.fit_hierarchical(radius_cutoff=r, member_cutoff=member_cutoff)
    assert canon_bundle(cl._bundle) == _canon_oracle(expect)
    assert (cl.labels == 1).all() and (cl._bundle._labels.consider == 0).all()


def test_mst_hierarchy_single_tree_root(cuda):
    # one dense blob: a spanning tree; the root takes the last merge's graph and lambda
    X = _blobs(200, 2, 9, k=1, spread=0.3)
    r = 0.6
    tree = O.mst_prim(O.neighbourhoods(X, r))
    assert len(tree) == len(X) - 1
    cl = Clustering(X, hfitter="mst")
    cl.fit_hierarchical(radius_cutoff=r, member_cutoff=5)
    expect = O.mst_hierarchy(len(X), tree, 5)
    expect.pop("top")
    assert canon_bundle(cl._bundle) == _canon_oracle(expect)


def _compare_repeat(bundle, node, path=()):
    if node["labels"] is None:
        assert bundle._labels is None, path
    else:
        np.testing.assert_array_equal(bundle.labels, node["labels"], err_msg=str(path))
    if node["root"] is None:
        assert bundle._reference_indices is None, path
    else:
        np.testing.assert_array_equal(bundle.root_indices, node["root"], err_msg=str(path))
        np.testing.assert_array_equal(bundle.parent_indices, node["parent"], err_msg=str(path))
    assert list(bundle.children) == list(node["children"]), path
    count = 1
    for k in node["children"]:
        count += _compare_repeat(bundle.children[k], node["children"][k], path + (k,))
    return count


@pytest.mark.parametrize("precision", ["f64", "f32_band"])
@pytest.mark.parametrize("radii,cuts", [([0.3], [1, 4, 9]),                 # one graph, growing c
                                        ([0.5, 0.3, 0.18], [3]),            # shrinking r: re-thresholded graphs
                                        ([0.4, 0.3, 0.3], [2, 6, 10])])     # r repeated, c varied
def test_repeat_hierarchy_matches_oracle(cuda, precision, radii, cuts):
    # steps that only shrink clusters: with a cluster spread over several parents
    # (r growing) the reference's own isolate fails with an IndexError
    X = _blobs(350, 2, 11)
    # the Repeat fitter passes r and c unchanged: reduced distances <= r, c_sim = m_core = c
    expect = O.hierarchical_repeat(X, radii, cuts, member_cutoff=4)
    fitter = FitterGPU(neighbours_getter=NeighboursGetterGPU(precision=precision))
    cl = Clustering(X, hierarchical_fitter=HierarchicalFitterRepeat(fitter))
    cl.fit_hierarchical(radius_cutoff=radii, cnn_cutoff=cuts, member_cutoff=4, v=False)
    assert _compare_repeat(cl._bundle, expect) > 3
    # parameters of the first step recorded per cluster label of the root
    assert all(v == (radii[0], cuts[0]) for v in cl._bundle._labels.meta["params"].values())


def test_repeat_hierarchy_unsorted_and_recipe(cuda, capsys):
    X = _blobs(300, 3, 12)
    expect = O.hierarchical_repeat(X, [0.6], [2, 5], sort_by_size=False)
    cl = Clustering(X, hfitter="repeat")
    cl.fit_hierarchical(radius_cutoff=0.6, cnn_cutoff=[2, 5], sort_by_size=False)
    _compare_repeat(cl._bundle, expect)
    out = capsys.readouterr().out
    assert "Running step 0     (r = 0.6, c = 2)" in out and "Running step 1" in out


def test_repeat_then_reel_and_trim(cuda):
    X = _blobs(400, 2, 13)
    cl = Clustering(X, hfitter="repeat")
    cl.fit_hierarchical(radius_cutoff=0.3, cnn_cutoff=[2, 8], member_cutoff=3, v=False)
    before = cl.labels.copy()
    cl.trim_trivial_leafs()
    cl.reel()
    after = cl.labels
    # reeling relabels points of clusters with children; noise of the first step stays noise
    assert ((before == 0) <= (after == 0)).all()
    assert cl._bundle._labels.meta["origin"] == "reel"


def _clusters_2d(seed, n=300):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-3, 3, size=(4, 2))
    return np.ascontiguousarray(centres[rng.integers(0, 4, n)] + rng.normal(0, 0.5, size=(n, 2)))


def test_repeat_hierarchy_precomputed_distances(cuda):
    # a matrix of reduced distances: D[i, j] <= r is the oracle's s_ij <= r
    P = _clusters_2d(21)
    D = O.reduced_distances(P, P)
    radii, cuts = [0.3, 0.2], [3, 6]
    expect = O.hierarchical_repeat(P, radii, cuts, member_cutoff=4)
    fitter = Clustering(D, recipe="distances").fitter
    cl = Clustering(D, recipe="distances", hierarchical_fitter=HierarchicalFitterRepeat(fitter))
    cl.fit_hierarchical(radius_cutoff=radii, cnn_cutoff=cuts, member_cutoff=4, v=False)
    assert _compare_repeat(cl._bundle, expect) > 3


def test_repeat_hierarchy_neighbourhood_input(cuda):
    P = _clusters_2d(22)
    r = 0.25
    lists = [np.nonzero(row <= r)[0] for row in O.reduced_distances(P, P)]
    expect = O.hierarchical_repeat(P, [r], [2, 5, 9], member_cutoff=3)
    fitter = Clustering(lists, recipe="neighbourhoods").fitter
    cl = Clustering(lists, recipe="neighbourhoods", hierarchical_fitter=HierarchicalFitterRepeat(fitter))
    cl.fit_hierarchical(radius_cutoff=r, cnn_cutoff=[2, 5, 9], member_cutoff=3, v=False)
    assert _compare_repeat(cl._bundle, expect) > 3


def test_repeat_hierarchy_periodic_metric(cuda):
    import os
    from commonnnclustering_amd._types import MetricExtEuclideanPeriodicReduced

    data = np.load(os.path.join(os.path.dirname(__file__), "golden", "backbone_dihedrals.npy"))[::3]
    data = np.ascontiguousarray(data)
    metric = (MetricExtEuclideanPeriodicReduced, (np.array([360.0, 360.0]),), {})
    fitter = Clustering(data, fitter__getter__dgetter__metric=metric).fitter
    cl = Clustering(data, hierarchical_fitter=HierarchicalFitterRepeat(fitter))
    # squared-distance thresholds (the Repeat fitter passes r unchanged), shrinking: graphs rebuilt per radius
    radii, cuts = [400.0, 150.0], [5, 10]
    cl.fit_hierarchical(radius_cutoff=radii, cnn_cutoff=cuts, member_cutoff=10, v=False)
    expect = O.hierarchical_repeat(data, radii, cuts, member_cutoff=10, bounds=[360.0, 360.0])
    assert _compare_repeat(cl._bundle, expect) > 2


def test_mst_hierarchy_neighbourhood_input(cuda):
    from commonnnclustering_amd._types import NeighboursGetterLookup

    P = _clusters_2d(23, 250)
    lists = [np.nonzero(row <= 0.2)[0] for row in O.reduced_distances(P, P)]
    tree = O.mst_prim(lists)
    expect = O.mst_hierarchy(len(P), tree, 5)
    expect.pop("top", None)
    cl = Clustering(lists, recipe="neighbourhoods",
                    hierarchical_fitter=HierarchicalFitterMSTPrim(neighbours_getter=NeighboursGetterLookup()))
    cl.fit_hierarchical(radius_cutoff=0.0, member_cutoff=5)
    assert canon_bundle(cl._bundle) == _canon_oracle(expect)
