"""Serving loop: every result is the same whether the request runs alone or
in a mix of concurrent requests (bit-identical labels), re-fits reuse the
resident graph, and the admission budget only changes the schedule."""

import numpy as np
import pytest
import torch

from oracle import c_oracle as C
from commonnnclustering_amd import engine, synthetic
from commonnnclustering_amd.serving import FitRequest, Server

pytestmark = pytest.mark.gpu


def _requests():
    reqs, data = [], {}
    for k in range(6):
        X, p = synthetic.serving_request(7000 + k, n_range=(1_000, 12_000), dims=(2, 3, 16, 64))
        name = f"ds{k}"
        data[name] = X
        reqs.append(FitRequest(p["radius_cutoff"], p["cnn_cutoff"], X=X, dataset_id=name,
                               member_cutoff=p["member_cutoff"]))
    # re-fits: smaller radius, other cutoff, both, and a larger radius (rebuild)
    base = {r.dataset_id: r for r in reqs}
    reqs.append(FitRequest(base["ds0"].radius_cutoff * 0.8, base["ds0"].cnn_cutoff, dataset_id="ds0"))
    reqs.append(FitRequest(base["ds1"].radius_cutoff, base["ds1"].cnn_cutoff + 3, dataset_id="ds1"))
    reqs.append(FitRequest(base["ds2"].radius_cutoff * 0.7, 2, dataset_id="ds2", member_cutoff=5))
    reqs.append(FitRequest(base["ds3"].radius_cutoff * 1.2, base["ds3"].cnn_cutoff, dataset_id="ds3"))
    return reqs, data


def _alone(req, data):
    X = torch.from_numpy(data[req.dataset_id]).cuda()
    r, c = req.radius_cutoff, req.cnn_cutoff
    lab, _, _, _ = engine.fit(X, r * r, c + 2, c + 1, precision=req.precision, member_cutoff=req.member_cutoff,
                              max_clusters=req.max_clusters)
    return lab.cpu().numpy()


def test_alone_equals_mixed(cuda):
    reqs, data = _requests()
    want = [_alone(r, data) for r in reqs]
    with Server(slots=4) as srv:
        futs = [srv.submit(r) for r in reqs]
        got = [f.result(timeout=120) for f in futs]
        summary = srv.latency_summary()
    for k, (g, w) in enumerate(zip(got, want)):
        assert (g.labels == w).all(), f"request {k} ({reqs[k].dataset_id})"
    assert [g.graph_reused for g in got[6:]] == [True, True, True, False]
    assert summary["requests"] == len(reqs) and summary["p99_ms"] >= summary["p50_ms"] > 0


def test_results_match_oracle(cuda):
    reqs, data = _requests()
    with Server(slots=3) as srv:
        futs = [srv.submit(r) for r in reqs]
        got = [f.result(timeout=120) for f in futs]
    for k in (0, 6, 8):
        r = reqs[k]
        want = C.fit(data[r.dataset_id], r.radius_cutoff, r.cnn_cutoff, member_cutoff=r.member_cutoff)
        assert (got[k].labels == want).all(), k


def test_tight_budget_serialises_but_agrees(cuda):
    reqs, data = _requests()
    want = [_alone(r, data) for r in reqs[:6]]
    with Server(slots=4, memory_budget=1 << 20, cache_budget=1 << 30) as srv:
        futs = [srv.submit(r) for r in reqs[:6]]
        got = [f.result(timeout=120) for f in futs]
    for g, w in zip(got, want):
        assert (g.labels == w).all()


def test_unknown_dataset_is_rejected(cuda):
    with Server(slots=1) as srv:
        with pytest.raises(KeyError):
            srv.submit(FitRequest(0.5, 3, dataset_id="missing"))


def test_evicted_dataset_refits_from_its_coordinates(cuda):
    """A cache too small to keep any graph: every re-fit finds its dataset
    evicted and runs as a fresh fit from the coordinates of the dataset's first
    request -- the same labels as a fit alone."""
    reqs, data = _requests()
    with Server(slots=2, cache_budget=1) as srv:
        first = [srv.submit(r) for r in reqs[:6]]
        for f in first:
            f.result(timeout=120)
        later = [srv.submit(r) for r in reqs[6:]]
        got = [f.result(timeout=120) for f in later]
    for k, g in enumerate(got):
        r = reqs[6 + k]
        assert (g.labels == _alone(r, data)).all(), r.dataset_id
        assert not g.graph_reused
