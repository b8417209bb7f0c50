"""Continuous-batching fit server on one GPU.

Fit requests of any n and d are queued, admitted as slots and device memory
free up, and run concurrently: every slot is a host thread with its own
library context and its own HIP stream (default priority), so the kernels of
admitted requests overlap on the device.  A request that names a dataset keeps
that dataset's neighbour graph resident; later requests for the same dataset
at a smaller or equal radius re-threshold it exactly (FP64) instead of
recomputing distances, and requests at another cnn_cutoff only rerun the
labelling pass.  Labels are those of a fresh ``Clustering.fit`` at the same
parameters (cluster.pyx:256-392) whatever else runs beside them.

    with Server(slots=4) as srv:
        fut = srv.submit(FitRequest(radius_cutoff=0.3, cnn_cutoff=5, X=points, dataset_id="a"))
        again = srv.submit(FitRequest(radius_cutoff=0.2, cnn_cutoff=8, dataset_id="a"))
        labels = fut.result().labels
    print(srv.latency_summary())
"""

from __future__ import annotations

import collections
import concurrent.futures as cf
import dataclasses
import itertools
import threading
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import engine


@dataclass
class FitRequest:
    """One fit: fresh (``X`` given) or a re-fit of the resident ``dataset_id``."""

    radius_cutoff: float
    cnn_cutoff: int
    X: Optional[object] = None          # numpy array or CUDA tensor (n, d), float32/float64
    dataset_id: Optional[str] = None    # keep / reuse the neighbour graph under this name
    member_cutoff: Optional[int] = None
    max_clusters: Optional[int] = None
    precision: str = "f32_band"


@dataclass
class FitResult:
    request_id: int
    dataset_id: Optional[str]
    labels: np.ndarray
    n_points: int
    n_clusters: int
    graph_reused: bool
    t_submit: float
    t_start: float
    t_end: float

    @property
    def latency_ms(self) -> float:
        return (self.t_end - self.t_submit) * 1e3

    @property
    def queue_ms(self) -> float:
        return (self.t_start - self.t_submit) * 1e3


@dataclass
class _Resident:
    graph: engine.NeighbourGraph
    X: torch.Tensor
    r2: float
    precision: str
    nbytes: int
    lock: threading.Lock = field(default_factory=threading.Lock)


@dataclass
class _Pending:
    rid: int
    req: FitRequest
    est: int
    future: cf.Future
    t_submit: float


def _estimate_bytes(n: int, d: int, r2: float, X, resident: Optional[_Resident]) -> int:
    """Device bytes a request may hold while it runs: coordinates and working
    copies, tile data, labels workspace and the neighbour lists, whose size is
    estimated from the fraction of sampled point pairs within the radius."""
    if resident is not None:
        return 2 * resident.nbytes // 3 + 64 * n
    base = n * d * 24 + 96 * n
    if n < 2:
        return base
    rng = np.random.default_rng(0)
    m = min(20000, n * (n - 1))
    i = rng.integers(0, n, size=m)
    j = rng.integers(0, n, size=m)
    if isinstance(X, torch.Tensor):
        Xs = X[torch.from_numpy(i).to(X.device)].double(), X[torch.from_numpy(j).to(X.device)].double()
        frac = float(((Xs[0] - Xs[1]) ** 2).sum(1).le(r2).double().mean())
    else:
        A, B = np.asarray(X[i], dtype=np.float64), np.asarray(X[j], dtype=np.float64)
        frac = float((((A - B) ** 2).sum(1) <= r2).mean())
    nnz = n + frac * n * n
    return int(base + 16 * nnz)


class Server:
    """Admits fit requests by device-memory budget into ``slots`` concurrent
    workers; keeps neighbour graphs of named datasets resident (LRU, bounded by
    ``cache_budget`` bytes)."""

    def __init__(self, device: int = 0, slots: int = 4, memory_budget: Optional[int] = None,
                 cache_budget: Optional[int] = None):
        self.device = torch.device("cuda", device)
        if memory_budget is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            memory_budget = int(0.6 * free)
        self.memory_budget = int(memory_budget)
        self.cache_budget = int(cache_budget if cache_budget is not None else self.memory_budget // 3)
        self._cv = threading.Condition()
        self._pending: collections.deque[_Pending] = collections.deque()
        self._inflight = 0
        self._running = 0
        self._closed = False
        self._cache: "collections.OrderedDict[str, _Resident]" = collections.OrderedDict()
        self._producing: dict = {}  # dataset_id -> fresh request that will make it resident
        # dataset_id -> the caller's coordinates of its latest fresh request (a
        # reference, no copy): a re-fit of a dataset whose graph was evicted runs
        # as a fresh fit from them
        self._sources: dict = {}
        self._ids = itertools.count()
        self.results: list[FitResult] = []
        self._workers = [threading.Thread(target=self._worker, name=f"cnn-slot-{k}", daemon=True)
                         for k in range(max(1, int(slots)))]
        for w in self._workers:
            w.start()

    # ------------------------------------------------------------------ API
    def submit(self, req: FitRequest) -> cf.Future:
        fut: cf.Future = cf.Future()
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            r2 = float(req.radius_cutoff) ** 2
            if req.X is None:
                # re-fit: of a resident dataset, or of one whose first fit is queued or running
                res = self._cache.get(req.dataset_id)
                prod = self._producing.get(req.dataset_id)
                if res is None and prod is None:
                    src = self._sources.get(req.dataset_id)
                    if src is None:
                        raise KeyError(f"no resident dataset {req.dataset_id!r}; send X with the first request")
                    # evicted: a fresh fit from the dataset's coordinates makes it resident again
                    req = dataclasses.replace(req, X=src)
                    self._producing[req.dataset_id] = req
                    est = _estimate_bytes(src.shape[0], src.shape[1], r2, src, None)
                elif res is not None and r2 <= res.r2:
                    est = _estimate_bytes(res.graph.n, res.graph.d, r2, None, res)
                else:
                    Xr = res.X if res is not None else prod.X
                    est = _estimate_bytes(Xr.shape[0], Xr.shape[1], r2, Xr, None)
            else:
                n, d = req.X.shape
                est = _estimate_bytes(n, d, r2, req.X, None)
                if req.dataset_id is not None:
                    self._producing[req.dataset_id] = req
                    self._sources[req.dataset_id] = req.X
            p = _Pending(next(self._ids), req, est, fut, time.perf_counter())
            self._pending.append(p)
            self._cv.notify_all()
        return fut

    def close(self, wait: bool = True) -> None:
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            for w in self._workers:
                w.join()
        with self._cv:
            for res in self._cache.values():
                res.graph.release()
            self._cache.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def resident_bytes(self) -> int:
        with self._cv:
            return sum(r.nbytes for r in self._cache.values())

    def latency_summary(self) -> dict:
        lat = np.array([r.latency_ms for r in self.results]) if self.results else np.zeros(0)
        if lat.size == 0:
            return {"requests": 0}
        t0 = min(r.t_submit for r in self.results)
        t1 = max(r.t_end for r in self.results)
        pts = sum(r.n_points for r in self.results)
        return {"requests": int(lat.size), "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                "points_per_s": pts / max(t1 - t0, 1e-9), "span_s": t1 - t0,
                "refits_reusing_graph": int(sum(r.graph_reused for r in self.results))}

    # ------------------------------------------------------------------ admission
    def _cache_bytes(self) -> int:
        return sum(r.nbytes for r in self._cache.values())

    def _evict(self, need: int) -> None:
        """Drop least recently used resident graphs (not in use) until `need`
        more bytes fit next to the running requests."""
        for key in list(self._cache.keys()):
            if self._inflight + self._cache_bytes() + need <= self.memory_budget:
                return
            res = self._cache[key]
            if res.lock.acquire(blocking=False):
                try:
                    res.graph.release()
                    del self._cache[key]
                finally:
                    res.lock.release()

    def _take(self) -> Optional[_Pending]:
        """Next admissible request: the first queued one whose estimate fits the
        free budget (an oversized request runs alone)."""
        with self._cv:
            while True:
                if self._closed and not self._pending:
                    return None
                for k, p in enumerate(self._pending):
                    if p.req.X is None and p.req.dataset_id not in self._cache and \
                            p.req.dataset_id in self._producing:
                        continue  # its dataset is not resident yet
                    if self._inflight + self._cache_bytes() + p.est > self.memory_budget:
                        self._evict(p.est)
                    fits = self._inflight + self._cache_bytes() + p.est <= self.memory_budget
                    if fits or self._running == 0:
                        del self._pending[k]
                        self._inflight += p.est
                        self._running += 1
                        return p
                self._cv.wait()

    def _done(self, p: _Pending) -> None:
        with self._cv:
            self._inflight -= p.est
            self._running -= 1
            if p.req.X is not None and self._producing.get(p.req.dataset_id) is p.req:
                del self._producing[p.req.dataset_id]
            self._cv.notify_all()

    # ------------------------------------------------------------------ work
    def _worker(self) -> None:
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(self.device)  # default priority
        with torch.cuda.stream(stream):
            while True:
                p = self._take()
                if p is None:
                    return
                t_start = time.perf_counter()
                try:
                    labels, reused, n = self._run(p.req)
                    res = FitResult(p.rid, p.req.dataset_id, labels, n, int(labels.max(initial=0)), reused,
                                    p.t_submit, t_start, time.perf_counter())
                    with self._cv:
                        self.results.append(res)
                    p.future.set_result(res)
                except BaseException as e:  # the future carries the error to the caller
                    p.future.set_exception(e)
                finally:
                    self._done(p)

    def _run(self, req: FitRequest):
        r2 = float(req.radius_cutoff) ** 2
        c = int(req.cnn_cutoff)
        if req.X is None:
            with self._cv:
                res = self._cache.get(req.dataset_id)
                if res is not None:
                    self._cache.move_to_end(req.dataset_id)
                src = self._sources.get(req.dataset_id)
            out = None if res is None else self._refit(res, req, r2, c)
            if out is not None:
                return out
            # evicted since it was queued: a fresh fit from the dataset's coordinates
            if src is None:
                raise KeyError(f"resident dataset {req.dataset_id!r} was evicted; send X again")
            req = dataclasses.replace(req, X=src)
        return self._fresh(req, r2, c)

    def _refit(self, res: _Resident, req: FitRequest, r2: float, c: int):
        """Re-fit of a resident dataset; None when its graph was released meanwhile."""
        with res.lock:
            if res.graph._handle is None:
                return None
            if r2 <= res.r2 and req.precision == res.precision:
                g = res.graph if r2 == res.r2 else res.graph.rethreshold(r2)
                try:
                    lab, _ = g.fit(c + 2, c + 1, sort=True, member_cutoff=req.member_cutoff,
                                   max_clusters=req.max_clusters)
                    out = lab.cpu().numpy()
                finally:
                    if g is not res.graph:
                        g.release()
                return out, True, res.graph.n
            # larger radius: a new graph replaces the resident one
            lab, g, _, _ = engine.fit(res.X, r2, c + 2, c + 1, precision=req.precision,
                                      member_cutoff=req.member_cutoff, max_clusters=req.max_clusters,
                                      keep_graph=True)
            out = lab.cpu().numpy()
            old = res.graph
            res.graph, res.r2, res.precision = g, r2, req.precision
            old.release()
            res.nbytes = g.nbytes + res.X.numel() * res.X.element_size()
            return out, False, g.n

    def _fresh(self, req: FitRequest, r2: float, c: int):
        X = req.X
        if not isinstance(X, torch.Tensor):
            X = torch.from_numpy(np.ascontiguousarray(X))
        X = X.to(self.device).contiguous()
        keep = req.dataset_id is not None
        lab, g, _, _ = engine.fit(X, r2, c + 2, c + 1, precision=req.precision, member_cutoff=req.member_cutoff,
                                  max_clusters=req.max_clusters, keep_graph=keep)
        out = lab.cpu().numpy()
        if keep:
            nbytes = g.nbytes + X.numel() * X.element_size()
            with self._cv:
                old = self._cache.pop(req.dataset_id, None)
                self._cache[req.dataset_id] = _Resident(g, X, r2, req.precision, nbytes)
                # keep the resident set within its own budget (least recently used first;
                # entries in use are skipped, never waited for while the queue lock is held)
                for key in list(self._cache.keys()):
                    if self._cache_bytes() <= self.cache_budget or key == req.dataset_id:
                        continue
                    res = self._cache[key]
                    if res.lock.acquire(blocking=False):
                        try:
                            res.graph.release()
                            del self._cache[key]
                        finally:
                            res.lock.release()
            if old is not None:  # a replaced entry may still be in use: wait outside the queue lock
                with old.lock:
                    old.graph.release()
        return out, False, X.shape[0]
