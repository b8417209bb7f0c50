"""cnn_labels_from_csr_noalloc (the labelling entry point that allocates no
device memory) against cnn_labels_from_csr, both called through the C-ABI:
same labels, consider flags and statistics on consecutive calls (the second
on a smaller input, so stale workspace contents would show), at the largest
accepted size, and clear errors beyond it."""

import ctypes

import numpy as np
import pytest
import torch

from commonnnclustering_amd import _lib
from commonnnclustering_amd._lib import LabelStats, ptr
from oracle import c_oracle as C

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 20


@pytest.fixture(scope="module")
def fresh_ctx(cuda):
    """A context of its own: its arena must stay empty across the calls."""
    h = ctypes.c_void_p()
    L = _lib.lib()
    assert L.cnn_ctx_create(0, ctypes.byref(h)) == 0
    yield h
    L.cnn_ctx_destroy(h)


def _blobs(n, seed):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-8, 8, size=(5, 2))
    lab = rng.integers(0, 5, size=n)
    X = centres[lab] + rng.normal(scale=0.6, size=(n, 2))
    return X


def _call(fn, ctx, rp, col, c_sim, m_core, consider, current_start, sim_mode, labels):
    st = LabelStats()
    rc = fn(ctx, ptr(rp), ptr(col), rp.numel() - 1, c_sim, m_core, ptr(consider), current_start, sim_mode,
            ptr(labels), ctypes.byref(st))
    return rc, st.to_dict()


def _both(fresh, rp_h, col_h, c_sim, m_core, seed, current_start=1, sim_mode=0):
    L = _lib.lib()
    ref_ctx = _lib.get_context(torch.device("cuda", 0))
    n = len(rp_h) - 1
    rp = torch.from_numpy(rp_h).cuda()
    col = torch.from_numpy(col_h if len(col_h) else np.zeros(1, np.int32)).cuda()
    rng = np.random.default_rng(seed)
    cons_h = (rng.random(n) < 0.9).astype(np.uint8)
    prior = rng.integers(0, 3, size=n).astype(np.int64)
    out = []
    for fn, ctx in ((L.cnn_labels_from_csr, ref_ctx.handle), (L.cnn_labels_from_csr_noalloc, fresh)):
        lab = torch.from_numpy(prior.copy()).cuda()
        cons = torch.from_numpy(cons_h.copy()).cuda()
        rc, st = _call(fn, ctx, rp, col, c_sim, m_core, cons, current_start, sim_mode, lab)
        assert rc == 0, L.cnn_last_error(ctx)
        torch.cuda.synchronize()
        out.append((lab.cpu().numpy(), cons.cpu().numpy(), st))
    (l1, c1, s1), (l2, c2, s2) = out
    assert np.array_equal(l1, l2)
    assert np.array_equal(c1, c2) and not c2.any()
    for k in ("n_core", "n_components", "n_clusters"):
        assert s1[k] == s2[k], k
    want, _ = C.labels_from_csr(rp_h, col_h, c_sim, m_core, consider=cons_h, current_start=current_start,
                                labels=prior.copy())
    assert np.array_equal(l2, want)
    return l2


def _arena_bytes(ctx):
    v = ctypes.c_int64(-1)
    assert _lib.lib().cnn_ctx_workspace_bytes(ctx, ctypes.byref(v)) == 0
    return v.value


def test_two_calls_second_smaller(fresh_ctx):
    X1 = _blobs(30000, 1)
    rp1, col1 = C.neighbours_csr(X1, 0.35)
    lab1 = _both(fresh_ctx, rp1, col1, 7, 6, seed=5, current_start=3)
    assert lab1.max() >= 4
    X2 = _blobs(2500, 2)
    rp2, col2 = C.neighbours_csr(X2, 0.5)
    _both(fresh_ctx, rp2, col2, 5, 4, seed=6)
    assert _arena_bytes(fresh_ctx) == 0


def test_largest_size(fresh_ctx):
    """n = CNN_NOALLOC_MAX_POINTS: rows i-2 .. i+2 (a chain), c_sim = 3."""
    n = NOALLOC_MAX
    off = np.arange(-2, 3)
    idx = np.arange(n)[:, None] + off[None, :]
    ok = (idx >= 0) & (idx < n)
    deg = ok.sum(axis=1)
    rp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(deg, out=rp[1:])
    col = idx[ok].astype(np.int32)
    _both(fresh_ctx, rp, col, 3, 2, seed=7)
    assert _arena_bytes(fresh_ctx) == 0


def test_rejects_beyond_limit_and_bitset(fresh_ctx):
    L = _lib.lib()
    n = NOALLOC_MAX + 1
    rp = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    col = torch.zeros(1, dtype=torch.int32, device="cuda")
    lab = torch.zeros(n, dtype=torch.int64, device="cuda")
    rc, _ = _call(L.cnn_labels_from_csr_noalloc, fresh_ctx, rp, col, 3, 2, None, 1, 0, lab)
    assert rc == _lib.CNN_ERR_INVALID
    assert b"no-allocation limit" in L.cnn_last_error(fresh_ctx)
    rp = torch.zeros(11, dtype=torch.int64, device="cuda")
    lab = torch.zeros(10, dtype=torch.int64, device="cuda")
    rc, _ = _call(L.cnn_labels_from_csr_noalloc, fresh_ctx, rp, col, 3, 2, None, 1, 2, lab)
    assert rc == _lib.CNN_ERR_INVALID
    assert _arena_bytes(fresh_ctx) == 0
