"""Autotuner: launch shapes and algorithm switches per problem class.

The library keeps a process-wide table of tuned shapes (``cnn_tune_set``,
include/cnn_gpu.h); every graph build and fit looks up the entry of its class
-- dimension bucket, size bucket (powers of two) and precision form -- and uses
the defaults when there is none.  No value changes a result, only the speed:
the tuner checks that every candidate gives the same labels.

Tuned here (``cnn_tune_params``):
  mfma_col_chunk  column tiles per workgroup of the dense MFMA distance kernel
                  (FP32-banded form, d >= 16)
  wide_col_chunk  column tiles per workgroup of the dense plain-FP64 kernel
  emit_waves      waves per workgroup of the neighbour-list emission
  root_passes     root-hooking passes before the labelling's full round
  words_min_fill  lists with at least this many entries per 64-bit word are
                  tested as per-row word lists (bitset), else merged

``tune(X, r, c, precision)`` searches one class by coordinate descent (each
parameter in turn over its candidates, the others at their best so far), on
the caller's data, and stores the winner in the table.  ``tune_workloads``
runs it over the synthetic workloads and writes the cache file that
``load_cache`` (called when the library is first loaded) reads back.

    python -m commonnnclustering_amd.autotune [--out PATH]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import _lib

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")

# dimension buckets (inclusive) and size buckets [lo, hi)
D_BUCKETS = [(1, 1), (2, 2), (3, 3), (4, 8), (9, 15), (16, 32), (33, 64), (65, 1 << 20)]
N_BUCKETS = [(0, 1 << 14), (1 << 14, 1 << 17), (1 << 17, 1 << 20), (1 << 20, 1 << 31)]

CANDIDATES = {
    "mfma_col_chunk": [32, 64, 128],
    "wide_col_chunk": [16, 32, 64],
    "emit_waves": [4, 8],
    "root_passes": [1, 2, 3],
    "words_min_fill": [4, 8, 1 << 30],
}


def bucket_of(d: int, n: int):
    db = next(b for b in D_BUCKETS if b[0] <= d <= b[1])
    nb = next(b for b in N_BUCKETS if b[0] <= n < b[1])
    return db, nb


def current(d: int, n: int, precision) -> dict:
    """The shapes a problem of this class uses now (table entry or defaults)."""
    t = _lib.TuneParams()
    rc = _lib.lib().cnn_tune_get(int(d), int(n), _lib.precision_code(precision), ctypes.byref(t))
    if rc != 0:
        raise _lib.CNNError(rc, "cnn_tune_get")
    return t.to_dict()


def set_entry(d_range, n_range, precision, params: dict) -> None:
    t = _lib.TuneParams(**{k: int(v) for k, v in params.items()})
    rc = _lib.lib().cnn_tune_set(int(d_range[0]), int(d_range[1]), int(n_range[0]), int(n_range[1]),
                                 _lib.precision_code(precision), ctypes.byref(t))
    if rc != 0:
        raise _lib.CNNError(rc, f"cnn_tune_set {params}")


def relevant(d: int, precision: str) -> list:
    """Parameters that can change the time of a fit of this class."""
    p = _lib.precision_code(precision)
    names = ["root_passes"]
    if d > 8:  # dense layout: lists are emitted, similarity form chosen by fill
        names += ["emit_waves", "words_min_fill"]
        names.append("mfma_col_chunk" if (p == _lib.CNN_F32_BAND and d >= 16) else "wide_col_chunk")
    else:  # spatial layout: low-degree graphs emit their lists
        names.append("emit_waves")
    return names


def _time_fit(Xd, r2, c, precision, member_cutoff, steps):
    import torch

    from . import engine

    labels = torch.empty(Xd.shape[0], dtype=torch.int64, device=Xd.device)
    engine.fit(Xd, r2, c + 2, c + 1, precision=precision, member_cutoff=member_cutoff, labels=labels)
    torch.cuda.synchronize()
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        engine.fit(Xd, r2, c + 2, c + 1, precision=precision, member_cutoff=member_cutoff, labels=labels)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)) * 1e3, labels.cpu().numpy()


def tune(X, radius_cutoff: float, cnn_cutoff: int, precision="f32_band", member_cutoff=None, steps: int = 3,
         log=None) -> dict:
    """Coordinate descent over the relevant parameters of X's class; the best
    shapes are stored in the table.  Returns the entry with its timings."""
    import torch

    from .engine import to_device_points

    Xd = to_device_points(X) if not isinstance(X, torch.Tensor) else X
    n, d = Xd.shape
    db, nb = bucket_of(d, n)
    r2 = float(radius_cutoff) ** 2
    best = current(d, n, precision)
    t_best, ref = _time_fit(Xd, r2, cnn_cutoff, precision, member_cutoff, steps)
    t_default = t_best
    trials = []
    for name in relevant(d, precision):
        for v in CANDIDATES[name]:
            if v == best[name]:
                continue
            cand = dict(best, **{name: v})
            set_entry(db, nb, precision, cand)
            t, lab = _time_fit(Xd, r2, cnn_cutoff, precision, member_cutoff, steps)
            if not np.array_equal(lab, ref):
                raise AssertionError(f"tuning changed the labels ({name}={v}): a shape must not change results")
            trials.append({"param": name, "value": v, "ms": t})
            if log:
                log(f"  {name}={v}: {t:.2f} ms (best {t_best:.2f})")
            if t < t_best:
                t_best, best = t, cand
        set_entry(db, nb, precision, best)
    return {"d_lo": db[0], "d_hi": db[1], "n_lo": nb[0], "n_hi": nb[1], "precision": precision, "params": best,
            "ms_default": t_default, "ms_tuned": t_best, "trials": trials, "n": int(n), "d": int(d)}


def _workloads():
    from . import synthetic

    out = []
    X, p = synthetic.md_basins_3d()
    out += [("md_basins_3d", X, p, "f32_band"), ("md_basins_3d", X, p, "f64")]
    X, p = synthetic.features_64d()
    out += [("features_64d", X, p, "f32_band"), ("features_64d", X, p, "f64")]
    X, p = synthetic.scan_2d()
    out.append(("scan_2d", X, {"radius_cutoff": float(p["radii"][-1]), "cnn_cutoff": 5, "member_cutoff": 2},
                "f32_band"))
    for seed, name in ((5003, "serving_d16"), (5001, "serving_d64")):
        X, p = synthetic.serving_request(seed, n_range=(40_000, 60_000), dims=(16,) if "16" in name else (64,))
        out.append((name, X, p, "f32_band"))
    for k, d in enumerate((2, 3, 8)):
        X, p = synthetic.one_small_dataset(4000 + k, n_range=(5_000, 10_000), dims=(d,))
        out.append((f"small_d{d}", X, p, "f32_band"))
    return out


def tune_workloads(path: str = DEFAULT_CACHE, steps: int = 3, log=print) -> dict:
    """Tunes every synthetic workload's class and writes the cache file."""
    import torch

    _lib.lib().cnn_tune_clear()
    entries = []
    for name, X, p, prec in _workloads():
        log(f"tuning {name} (n={X.shape[0]}, d={X.shape[1]}, {prec})")
        e = tune(X, p["radius_cutoff"], p["cnn_cutoff"], prec, member_cutoff=p.get("member_cutoff"), steps=steps,
                 log=log)
        e["workload"] = name
        entries.append(e)
        log(f"  -> {e['params']}: {e['ms_default']:.2f} -> {e['ms_tuned']:.2f} ms")
    doc = {"device": torch.cuda.get_device_name(0), "command": "python -m commonnnclustering_amd.autotune",
           "entries": entries}
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
    return doc


def load_cache(path: str = DEFAULT_CACHE) -> int:
    """Stores every entry of a cache file in the table; returns their number."""
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        doc = json.load(f)
    k = 0
    for e in doc.get("entries", []):
        set_entry((e["d_lo"], e["d_hi"]), (e["n_lo"], e["n_hi"]), e["precision"], e["params"])
        k += 1
    return k


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=DEFAULT_CACHE)
    ap.add_argument("--steps", type=int, default=3)
    args = ap.parse_args(argv)
    tune_workloads(args.out, args.steps, log=lambda m: print(m, file=sys.stderr, flush=True))


if __name__ == "__main__":
    main()
