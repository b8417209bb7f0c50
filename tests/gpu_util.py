"""Helpers for GPU parity tests: run the engine, bring results to the host."""

import numpy as np
import torch

from commonnnclustering_amd import engine


def to_dev(X):
    return torch.from_numpy(np.ascontiguousarray(X)).cuda()


def gpu_csr(X, r, precision="f64"):
    Xd = to_dev(X)
    r = float(r)
    row_ptr, col, st = engine.radius_csr(Xd, r * r, precision)
    return row_ptr.cpu().numpy(), col.cpu().numpy().astype(np.int64), st


def gpu_fit(X, r, c, member_cutoff=None, max_clusters=None, precision="f64", sort=True):
    Xd = to_dev(X)
    r = float(r)
    lab, _, gs, ls = engine.fit(Xd, r * r, c + 2, c + 1, precision=precision, sort=sort,
                                member_cutoff=member_cutoff, max_clusters=max_clusters)
    return lab.cpu().numpy(), gs, ls
