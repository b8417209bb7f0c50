"""Benchmark: CommonNN fit throughput on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W]
                [--workload md_basins_3d|features_64d|blobs_2d|serving|scan_2d|many_small|
                            hierarchy_repeat|hierarchy_mst]

A step is one full fit (radius graph, similarity test, union-find,
sort_by_size) of the whole synthetic dataset; the coordinates are already
resident in HBM when the timed region starts.  The timed K steps use the
FP32-banded form (``value``); the plain FP64 form is timed the same way and
reported beside it, and the two label vectors must be identical.

``roofline`` describes the longest single kernel of the fit (the distance
kernel, VALU- or MFMA-bound; the CSR emit kernel when a fit writes lists; the
label pass's full round, HBM-bound); ``kernel_rooflines`` and
``similarity_rooflines`` hold the others.  Durations come from HIP events the
library records around each launch on the stream the kernel runs on, averaged
over the timed fits.  ``cpu_baseline`` times the C restatement of the
reference fit (brute-force FP64 neighbours, components, sort_by_size) on all
host threads over a seeded subsample, extrapolated to the full workload by the
brute force's n^2 cost, and calibrated against the reference's published
single-thread tutorial timings.

The default line (md_basins_3d) also carries ``features_64d`` (n=2e5, d=64:
FP32-banded MFMA form against plain FP64, their ratio and the MFMA roofline)
and ``served`` (the serving workload: capacity and p50/p99 latency at Poisson
arrival rates of 0.5, 0.7 and 0.9 of capacity).

One JSON line is printed on stdout; progress goes to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fit points/s (plain FP64 vs FP32-banded) bit-exact labels; served p50/p99 ms"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3     # vector and dense MFMA FP32
FP64_PEAK_TFLOPS = 78.6      # FMA-counted; separately rounded sub/mul/add reach half of it

# HBM traffic per launch, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs
# scaled by a copy kernel of known size (profiles/*_traffic.json).
TRAFFIC_FILES = {
    ("k_dist_f32_cols", "md_basins_3d"): "profiles/k_dist_f32_cols_md_basins_3d_traffic.json",
    ("k_dist_mfma_lds", "features_64d"): "profiles/k_dist_mfma_lds_features_64d_traffic.json",
    ("round", "md_basins_3d"): "profiles/round_md_basins_3d_traffic.json",
    ("hook", "md_basins_3d"): "profiles/hook_md_basins_3d_traffic.json",
    ("round", "features_64d"): "profiles/round_features_64d_traffic.json",
    ("hook", "features_64d"): "profiles/hook_features_64d_traffic.json",
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name):
    from commonnnclustering_amd import synthetic

    gens = {"md_basins_3d": synthetic.md_basins_3d, "features_64d": synthetic.features_64d,
            "blobs_2d": synthetic.blobs_2d}
    if name not in gens:
        raise SystemExit(f"unknown workload {name}; choose from {sorted(gens)}")
    return gens[name]()


def run_fits(Xd, p, precision, steps, warmup):
    import torch
    from commonnnclustering_amd import engine

    r2 = float(p["radius_cutoff"]) ** 2
    c = int(p["cnn_cutoff"])
    m = p["member_cutoff"]
    labels = torch.empty(Xd.shape[0], dtype=torch.int64, device=Xd.device)
    for _ in range(warmup):
        engine.fit(Xd, r2, c + 2, c + 1, precision=precision, member_cutoff=m, labels=labels)
    torch.cuda.synchronize()
    gstats, lstats = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        _, _, gs, ls = engine.fit(Xd, r2, c + 2, c + 1, precision=precision, member_cutoff=m, labels=labels)
        gstats.append(gs)
        lstats.append(ls)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) / steps, labels.cpu().numpy(), gstats, lstats


def mean(stats, key):
    return float(np.mean([s[key] for s in stats]))


def phases(gs, ls):
    return {"prep": mean(gs, "ms_prep"), "distance": mean(gs, "ms_distance"), "emit": mean(gs, "ms_emit"),
            "similarity_unionfind": mean(ls, "ms_similarity"), "labels_sort": mean(ls, "ms_labels")}


def emit_roofline(gs, n, workload_name):
    """CSR emit kernel: compulsory HBM bytes per launch are the neighbour
    lists written (4 B per entry) and the row offsets read (8 B per row)."""
    launches = max(mean(gs, "distance_launches"), 1.0)
    ms = mean(gs, "ms_emit") / launches
    nbytes = (4.0 * mean(gs, "nnz") + 8.0 * (n + 1)) / launches
    achieved = nbytes / (ms * 1e-3) / 1e9
    traffic = None
    path = TRAFFIC_FILES.get(("k_emit_rows", workload_name))
    if path and os.path.exists(os.path.join(ROOT, path)):
        with open(os.path.join(ROOT, path)) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    return {"kernel": "k_emit_rows", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "ms_per_launch": ms,
            "algorithmic_bytes_per_launch": nbytes,
            "timing": "hipEvent pair around the launch on the library stream"}


def distance_roofline(gs, d, precision, workload_name):
    """Algorithmic work of the distance kernel per launch: 3 d flops (subtract,
    multiply, add) per evaluated point pair for the direct forms, 2 d for the
    Gram form on MFMA; pairs = point pairs the kernel computed (64x64 tile
    pairs, less the row groups of a pair out of reach in the spatial layout;
    symmetric evaluation computes each unordered tile pair once)."""
    launches = max(mean(gs, "distance_launches"), 1.0)
    # the main kernel alone (the distance phase also holds guard-band settling)
    ms = mean(gs, "ms_distance_kernel") / launches
    pairs = mean(gs, "pairs_evaluated") / launches
    gram = precision == "f32_band" and d >= 16
    # direct forms: subtract, multiply, add per dimension; Gram form: one FMA (2 flops) per dimension
    flops = (2.0 if gram else 3.0) * d * pairs
    if precision == "f32_band":
        if gram:
            kernel = "k_dist_mfma_lds" if (d + 2 + 7) // 8 <= 10 else "k_dist_mfma"
        else:
            kernel = "k_dist_f32_cols" if d <= 8 else "k_dist_f32_generic"
        bound, peak = ("mfma" if gram else "valu"), FP32_PEAK_TFLOPS
    else:
        kernel = "k_dist_f64_cols" if d <= 8 else "k_dist_f64_wide"
        bound, peak = "valu", FP64_PEAK_TFLOPS
    achieved = flops / (ms * 1e-3) / 1e12
    traffic = None
    path = TRAFFIC_FILES.get((kernel, workload_name))
    if path and os.path.exists(os.path.join(ROOT, path)):
        with open(os.path.join(ROOT, path)) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    return {"kernel": kernel, "bound": bound, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic, "ms_per_launch": ms, "flops_per_launch": flops,
            "pairs_per_launch": pairs, "timing": "hipEvent pair around the launch on the library stream"}


# Published single-thread fit times of the reference (cnnclustering 0.5.0,
# Cython, brute-force neighbours) for the six n=2000 datasets of its
# scikit-learn tutorial (docsrc/tutorial/scikit_learn_datasets.ipynb:384-414).
TUTORIAL_FITS = [  # dataset, radius_cutoff, cnn_cutoff, member_cutoff, published seconds
    ("circles", 0.5, 20, 100, 0.173), ("moons", 0.5, 20, 2, 0.193), ("varied", 0.28, 20, 20, 0.254),
    ("aniso", 0.29, 30, 5, 0.160), ("blobs", 0.4, 20, 2, 0.247), ("None", 0.5, 20, 1, 0.124),
]


def cpu_baseline(X, p, target_s=15.0, seed=0):
    """C restatement of the reference fit on all host threads, brute-force FP64
    neighbours as NeighboursGetterExtBruteForce (cost ~ n^2): a seeded
    subsample sized for about target_s seconds; points/s is extrapolated to the
    full workload by the n^2 cost.

    Calibration against the reference itself: the restatement is timed on one
    thread on the six tutorial datasets whose single-thread reference fit times
    are published (another machine), and on one thread on the first sample,
    which gives its multi-thread speed-up; the reference's rate on this host is
    estimated as value / speed-up / (published time / restatement time)."""
    from commonnnclustering_amd import synthetic
    from oracle import c_oracle

    n = X.shape[0]
    rng = np.random.default_rng(seed)
    cores = c_oracle.num_threads()

    def sample(m):
        idx = np.sort(rng.choice(n, size=min(m, n), replace=False))
        return np.ascontiguousarray(X[idx], dtype=np.float64)

    def run(Xs):
        t0 = time.perf_counter()
        c_oracle.fit(Xs, p["radius_cutoff"], p["cnn_cutoff"], member_cutoff=p["member_cutoff"], use_grid=False)
        return time.perf_counter() - t0

    X0 = sample(8000)
    m0, t0 = len(X0), run(X0)
    m = min(n, int(m0 * np.sqrt(target_s / max(t0, 1e-3))))
    if m > m0:
        Xm = sample(m)
        ms, ts = len(Xm), run(Xm)
    else:
        ms, ts = m0, t0
    t_full = ts * (n / ms) ** 2
    value = n / t_full
    # calibration: one thread
    tut = synthetic.sklearn_tutorial()
    c_oracle.set_num_threads(1)
    try:
        t1_sample = run(X0)
        ratios, port_s = [], []
        for name, r, c, mc, pub in TUTORIAL_FITS:
            Xt = np.ascontiguousarray(tut[name], dtype=np.float64)
            t_a = time.perf_counter()
            for _ in range(3):
                c_oracle.fit(Xt, r, c, member_cutoff=mc, use_grid=False)
            t_port = (time.perf_counter() - t_a) / 3
            port_s.append(t_port)
            ratios.append(pub / t_port)
    finally:
        c_oracle.set_num_threads(cores)
    speedup = t1_sample / max(t0, 1e-9)
    factor = float(np.median(ratios))
    return {"value": value, "unit": "points/s", "cores": cores, "kind": "port",
            "sample": f"{ms} of {n} points (seed {seed}), brute-force FP64 neighbour lists + components + "
                      f"sort_by_size in {ts:.2f} s ({ms / ts:.0f} points/s on the sample); full-size time "
                      f"extrapolated by (n/{ms})^2 to {t_full:.0f} s",
            "calibration": {
                "published_reference_s": [f[4] for f in TUTORIAL_FITS],
                "port_one_thread_s": port_s,
                "reference_over_port": factor,
                "port_speedup_on_cores": speedup,
                "reference_estimate_points_per_s": value / speedup / factor,
                "note": "six n=2000 tutorial fits (scikit_learn_datasets.ipynb:384-414): published single-thread "
                        "reference times (another machine) over the restatement's one-thread times on this "
                        "host; the reference estimate is value / speed-up / that factor"}}


def serving_requests(count, seed=5000, refit_frac=0.3):
    """Requests of the serving workload: fresh fits of synthetic datasets
    (n log-uniform in [1e3, 3e5], d in {2, 3, 16, 64}) and, for about
    refit_frac of them, re-fits of an earlier dataset at a smaller radius or
    another cnn_cutoff."""
    from commonnnclustering_amd import synthetic
    from commonnnclustering_amd.serving import FitRequest

    rng = np.random.default_rng(seed)
    reqs, fresh = [], []
    for k in range(count):
        if fresh and rng.uniform() < refit_frac:
            base = fresh[int(rng.integers(0, len(fresh)))]
            if rng.uniform() < 0.5:
                reqs.append(FitRequest(base.radius_cutoff * float(rng.uniform(0.6, 1.0)), base.cnn_cutoff,
                                       dataset_id=base.dataset_id, member_cutoff=base.member_cutoff))
            else:
                reqs.append(FitRequest(base.radius_cutoff, int(rng.integers(2, 51)), dataset_id=base.dataset_id,
                                       member_cutoff=base.member_cutoff))
            continue
        X, p = synthetic.serving_request(seed + k)
        r = FitRequest(p["radius_cutoff"], p["cnn_cutoff"], X=X, dataset_id=f"req{k}",
                       member_cutoff=p["member_cutoff"])
        reqs.append(r)
        fresh.append(r)
    return reqs


def su_rooflines(ls, workload_name):
    """Similarity + union-find kernels.  achieved = HBM bytes per launch
    (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated by a copy kernel) over the
    kernel's average duration (HIP events on the library stream).  Context: the
    phase's canonical algorithmic bytes B_S (every candidate edge merged with
    early exit at c + 2) + B_U, computed once by oracle/similarity_work.py; the
    union-find tests only a small share of the candidate edges, so B over the
    phase time can exceed the HBM peak -- that effective rate is not a frac."""
    out = {}
    # the full round: k_round_tile + k_round_queue on tile-pair words (k_round_b
    # on other graphs); the hook phase: k_hook_tile + k_hook_queue on row words,
    # k_hook on lists
    for kernel, key in (("round", "ms_round"), ("hook", "ms_hook")):
        ms = mean(ls, key)
        traffic = None
        path = TRAFFIC_FILES.get((kernel, workload_name))
        if path and os.path.exists(os.path.join(ROOT, path)):
            with open(os.path.join(ROOT, path)) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        achieved = traffic / (ms * 1e-3) / 1e9 if (traffic and ms > 0) else None
        out[kernel] = {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                       "ms_per_launch": ms,
                       "achieved_from": "measured HBM bytes per launch / kernel time (hipEvent pair)"}
    phase_ms = mean(ls, "ms_similarity")
    path = os.path.join(ROOT, "profiles", f"similarity_work_{workload_name}.json")
    if os.path.exists(path):
        with open(path) as f:
            w = json.load(f)
        b = w["B_S_bytes"] + w["B_U_bytes"]
        out["phase"] = {"ms": phase_ms, "B_S_bytes": w["B_S_bytes"], "B_U_bytes": w["B_U_bytes"],
                        "candidate_edges": w["candidate_edges"], "tests_executed": mean(ls, "tests"),
                        "effective_GBs": b / (phase_ms * 1e-3) / 1e9,
                        "note": "canonical work of the whole phase over its time; the union-find skips most "
                                "candidate edges, so this is not bounded by the HBM peak"}
    return out


def served_block(args, count=None, fracs=(0.5, 0.7, 0.9)):
    """Serving workload: capacity from a closed loop (every request queued at
    once), then Poisson arrivals at each fraction of that capacity; every rate
    replays the same requests on a fresh server (cold graph cache)."""
    from commonnnclustering_amd.serving import Server

    count = count or args.requests
    reqs = serving_requests(count)
    n_pts = [r.X.shape[0] for r in reqs if r.X is not None]
    log(f"serving: {len(reqs)} requests, {sum(r.X is None for r in reqs)} re-fits, "
        f"n in [{min(n_pts)}, {max(n_pts)}]")
    # warm-up: small requests page in code and pools
    with Server(slots=args.slots) as srv:
        for r in reqs[:6]:
            if r.X is not None:
                srv.submit(type(r)(r.radius_cutoff, r.cnn_cutoff, X=r.X[:2000])).result()
    with Server(slots=args.slots) as srv:
        t0 = time.perf_counter()
        futs = [srv.submit(r) for r in reqs]
        for f in futs:
            f.result()
        t_closed = time.perf_counter() - t0
        closed = srv.latency_summary()
    cap = len(reqs) / t_closed
    log(f"closed loop: {t_closed * 1e3:.1f} ms, capacity {cap:.1f} req/s, p50 {closed['p50_ms']:.1f} ms, "
        f"p99 {closed['p99_ms']:.1f} ms")
    sweep = []
    for k, frac in enumerate(fracs):
        rate = frac * cap
        gaps = np.random.default_rng(1 + k).exponential(1.0 / rate, size=len(reqs))
        with Server(slots=args.slots) as srv:
            futs = []
            t_next = time.perf_counter()
            for r, g in zip(reqs, gaps):
                t_next += g
                dt = t_next - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                futs.append(srv.submit(r))
            for f in futs:
                f.result()
            summ = srv.latency_summary()
        log(f"poisson {frac:.1f} x capacity ({rate:.1f} req/s): p50 {summ['p50_ms']:.1f} ms, "
            f"p99 {summ['p99_ms']:.1f} ms, {summ['points_per_s'] / 1e6:.1f} M points/s")
        sweep.append(dict(summ, rate_frac=frac, rate_req_per_s=rate))
    return {"requests_per_rate": len(reqs), "slots": args.slots, "refit_frac": 0.3,
            "n_range": [1000, 300000], "dims": [2, 3, 16, 64], "capacity_req_per_s": cap,
            "closed_loop": closed, "poisson": sweep}


def serving_bench(args):
    served = served_block(args)
    closed = served["closed_loop"]
    return {
        "metric": METRIC,
        "value": closed["points_per_s"],
        "unit": "points/s",
        "n_gpus": 1,
        "steps": served["requests_per_rate"],
        "warmup": 1,
        "ms_per_step": 1e3 / served["capacity_req_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded numpy generator)",
        "config": {"workload": "serving", "requests": served["requests_per_rate"], "slots": args.slots,
                   "n_range": [1000, 300000], "dims": [2, 3, 16, 64], "refit_frac": 0.3},
        "served": served,
    }


def neighbour_mismatches(Xd, r2):
    """Rows whose neighbour sets differ between the FP32-banded and the plain
    FP64 graph (the plain form is the reference arithmetic), entries compared."""
    import torch
    from commonnnclustering_amd import engine

    out = []
    for prec in ("f32_band", "f64"):
        g = engine.build_graph(Xd, r2, prec)
        try:
            out.append(g.to_csr())
        finally:
            g.release()
    (rp_a, col_a), (rp_b, col_b) = out
    deg_a, deg_b = rp_a[1:] - rp_a[:-1], rp_b[1:] - rp_b[:-1]
    bad = deg_a != deg_b
    if not bool(bad.any()):
        diff = (col_a != col_b).nonzero().flatten()
        if diff.numel():
            rows = torch.bucketize(diff, rp_a[1:], right=True)
            bad[rows] = True
    return int(bad.sum()), int(col_b.numel())


def features_block(args):
    """features_64d (n=2e5, d=64, float32): FP32-banded (MFMA Gram form with the
    FP64 guard band) against the plain FP64 form, per fit."""
    import torch
    from commonnnclustering_amd import autotune

    X, p = workload("features_64d")
    n, d = X.shape
    Xd = torch.from_numpy(X).cuda()
    steps = max(2, args.steps)
    t_band, lab_band, gs_band, ls_band = run_fits(Xd, p, "f32_band", steps, max(1, args.warmup))
    t_plain, lab_plain, gs_plain, ls_plain = run_fits(Xd, p, "f64", max(2, steps // 2), 1)
    rows_bad, entries = neighbour_mismatches(Xd, float(p["radius_cutoff"]) ** 2)
    log(f"features_64d: f32_band {t_band * 1e3:.2f} ms/fit, plain f64 {t_plain * 1e3:.2f} ms/fit "
        f"({t_plain / t_band:.2f}x); neighbour rows differing {rows_bad}")
    return {"config": {"workload": "features_64d", "n": int(n), "d": int(d), "dtype_in": str(X.dtype),
                       "radius_cutoff": float(p["radius_cutoff"]), "cnn_cutoff": int(p["cnn_cutoff"]),
                       "member_cutoff": int(p["member_cutoff"])},
            "f32_band": {"points_per_s": n / t_band, "ms_per_step": t_band * 1e3, "steps": steps},
            "plain_f64": {"points_per_s": n / t_plain, "ms_per_step": t_plain * 1e3,
                          "distance_roofline": distance_roofline(gs_plain, d, "f64", "features_64d")},
            "band_vs_plain": t_plain / t_band,
            "labels_identical": bool(np.array_equal(lab_band, lab_plain)),
            "neighbour_rows_differing_from_plain_f64": rows_bad, "neighbour_entries": entries,
            "band_hits": int(gs_band[-1]["band_hits"]), "band_flips": int(gs_band[-1]["band_flips"]),
            "max_fp32_abs_err_band": float(gs_band[-1]["max_abs_err_band"]),
            "phase_ms": phases(gs_band, ls_band), "phase_ms_plain": phases(gs_plain, ls_plain),
            "label_stats": ls_band[-1],
            "tuning": {"f32_band": autotune.current(d, n, "f32_band"), "f64": autotune.current(d, n, "f64")},
            "mfma_roofline": distance_roofline(gs_band, d, "f32_band", "features_64d"),
            "similarity_rooflines": su_rooflines(ls_band, "features_64d")}


def _timed(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / max(steps, 1)


def _line(value, ms, steps, warmup, config, extra):
    out = {"metric": METRIC, "value": value, "unit": "points/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64", "data": "synthetic (seeded numpy generator)", "config": config}
    out.update(extra)
    return out


def scan_bench(args):
    """(r, c) scan of one dataset (n=5e4, d=2, 8 radii x 8 cutoffs): one graph at
    the largest radius, exact FP64 re-thresholding, a labelling pass per pair.
    A step is the whole scan; points/s counts n per (r, c) pair.  Beside it:
    the same 64 labels from independent fits."""
    import torch
    from commonnnclustering_amd import engine, synthetic

    X, p = synthetic.scan_2d()
    Xd = torch.from_numpy(X).cuda()
    n = X.shape[0]
    radii, cs = list(p["radii"]), list(p["cnn_cutoffs"])
    pairs = len(radii) * len(cs)
    t = _timed(lambda: engine.fit_scan(Xd, radii, cs, precision="f32_band"), args.steps, args.warmup)

    def independent():
        for r in radii:
            for c in cs:
                engine.fit(Xd, float(r) * float(r), c + 2, c + 1, precision="f32_band")

    t_ind = _timed(independent, 1, 1)
    # the scan recorded once into a hipGraph on a resident graph, then replayed
    g = engine.build_graph(Xd, float(max(radii)) ** 2, "f32_band")
    t0 = time.perf_counter()
    sc = g.scan(radii, cs)
    t_capture = time.perf_counter() - t0
    t_replay = _timed(sc.launch, args.steps, 1)
    sc.release()
    g.release()
    log(f"scan: {t * 1e3:.2f} ms for {pairs} (r, c) pairs (graph build, capture, one launch); replay of the "
        f"captured scan {t_replay * 1e3:.2f} ms (capture {t_capture * 1e3:.1f} ms); independent fits "
        f"{t_ind * 1e3:.2f} ms")
    return _line(n * pairs / t, t * 1e3, args.steps, args.warmup,
                 {"workload": "scan_2d", "n": n, "d": 2, "radii": [float(r) for r in radii], "cnn_cutoffs": cs},
                 {"independent_fits_ms": t_ind * 1e3, "scan_vs_independent": t_ind / t,
                  "captured_replay_ms": t_replay * 1e3, "capture_ms": t_capture * 1e3,
                  "replay_points_per_s": n * pairs / t_replay})


def batch_bench(args):
    """512 small datasets (n log-uniform in [500, 2e4], d in {2, 3, 8}, r at the
    2nd distance percentile, c = 5), one cnn_fit_batch call per dimension.
    A step fits all 512; points/s counts every point once.  Beside it: the
    same datasets fitted one call after another."""
    import torch
    from commonnnclustering_amd import engine, synthetic

    sets = synthetic.many_small()
    groups = []
    for d in sorted({x.shape[1] for x, _ in sets}):
        mine = [(x, q) for x, q in sets if x.shape[1] == d]
        X = np.concatenate([x for x, _ in mine]).astype(np.float32)
        off = np.zeros(len(mine) + 1, dtype=np.int64)
        off[1:] = np.cumsum([len(x) for x, _ in mine])
        r2 = np.array([float(q["radius_cutoff"]) ** 2 for _, q in mine])
        c = np.array([int(q["cnn_cutoff"]) for _, q in mine])
        groups.append((torch.from_numpy(X).cuda(), off, r2, c))
    total = sum(len(x) for x, _ in sets)

    def batched():
        for Xd, off, r2, c in groups:
            engine.fit_batch(Xd, off, r2, c + 2, c + 1, precision="f32_band", member_cutoff=2)

    singles = [(torch.from_numpy(x.astype(np.float32)).cuda(), q) for x, q in sets]

    def one_by_one():
        for Xd, q in singles:
            r = float(q["radius_cutoff"])
            engine.fit(Xd, r * r, q["cnn_cutoff"] + 2, q["cnn_cutoff"] + 1, precision="f32_band", member_cutoff=2)

    t = _timed(batched, args.steps, args.warmup)
    t_seq = _timed(one_by_one, 1, 1)
    log(f"batch: {len(sets)} datasets, {total} points in {t * 1e3:.1f} ms; one call each {t_seq * 1e3:.1f} ms")
    return _line(total / t, t * 1e3, args.steps, args.warmup,
                 {"workload": "many_small", "datasets": len(sets), "points": total, "dims": [2, 3, 8]},
                 {"one_call_each_ms": t_seq * 1e3, "batch_vs_one_call_each": t_seq / t})


def hierarchy_repeat_bench(args):
    """Repeat hierarchy (HierarchicalFitterRepeat) of the md_basins_3d points
    (n=1e6, d=3): four steps at one radius (reduced distance 0.0625, the fit's
    graph) with cnn_cutoff 5, 10, 20, 40, member_cutoff 50, FP32-banded graph.
    A step is one fit_hierarchical call: step fits on one resident graph plus
    the host bundle tree (isolate per parent).  Beside it: the CPU restatement's
    fit rate (as the md_basins_3d baseline) divided by the number of steps."""
    import torch
    from commonnnclustering_amd import Clustering, synthetic
    from commonnnclustering_amd._fit import FitterGPU, HierarchicalFitterRepeat
    from commonnnclustering_amd._types import NeighboursGetterGPU

    X, _ = synthetic.md_basins_3d()
    n = X.shape[0]
    Xd = torch.from_numpy(X).cuda()
    cuts = [5, 10, 20, 40]
    fitter = FitterGPU(neighbours_getter=NeighboursGetterGPU(precision="f32_band"))
    hf = HierarchicalFitterRepeat(fitter)
    cl = Clustering(Xd, hierarchical_fitter=hf)

    def run():
        cl.fit_hierarchical(radius_cutoff=0.0625, cnn_cutoff=cuts, member_cutoff=50, v=False)

    t = _timed(run, args.steps, args.warmup)
    params = [hf.make_parameters(0.0625, c, 1) for c in cuts]
    t_steps = _timed(lambda: fitter.fit_steps(cl._bundle._input_data, params, True, 50, None), args.steps, 1)

    def count(b):
        return 1 + sum(count(c) for c in b.children.values())

    nodes = count(cl._bundle)
    cpu = None
    if not args.no_cpu:
        # the reference runs one full fit per step: the C fit restatement's rate over the steps
        one = cpu_baseline(X, {"radius_cutoff": 0.25, "cnn_cutoff": 10, "member_cutoff": 50},
                           target_s=args.cpu_seconds)
        cpu = dict(one, value=one["value"] / len(cuts),
                   sample=f"one fit per step ({len(cuts)} steps) at the C restatement's fit rate: " + one["sample"])
    log(f"hierarchy_repeat: {t * 1e3:.1f} ms per hierarchy ({len(cuts)} steps, {nodes} bundles); "
        f"step fits {t_steps * 1e3:.1f} ms")
    return _line(n / t, t * 1e3, args.steps, args.warmup,
                 {"workload": "hierarchy_repeat", "n": n, "d": 3, "radius_cutoff": 0.0625, "cnn_cutoff": cuts,
                  "member_cutoff": 50},
                 {"step_fits_ms": t_steps * 1e3, "host_tree_ms": (t - t_steps) * 1e3, "bundles": nodes,
                  "cpu_baseline": cpu})


def hierarchy_mst_bench(args):
    """MST hierarchy (HierarchicalFitterMSTPrim) of 1e5 seeded points of the
    md_basins_3d data (d=3) at radius_cutoff 0.25, member_cutoff 50.  A step is
    one fit_hierarchical call: radius graph, per-entry common-neighbour
    counts (k_edge_common), the host Prim forest and the bundle hierarchy.
    The weights call (CSR check + kernel) is also timed alone on the host clock."""
    import torch
    from commonnnclustering_amd import Clustering, engine, synthetic
    from commonnnclustering_amd._fit import HierarchicalFitterMSTPrim
    from commonnnclustering_amd._types import NeighboursGetterGPU

    X, _ = synthetic.md_basins_3d()
    rng = np.random.default_rng(1)
    m = 100_000
    Xs = np.ascontiguousarray(X[rng.choice(X.shape[0], m, replace=False)])
    Xd = torch.from_numpy(Xs).cuda()
    hf = HierarchicalFitterMSTPrim(neighbours_getter=NeighboursGetterGPU(precision="f32_band"))
    cl = Clustering(Xd, hierarchical_fitter=hf)

    def run():
        cl.fit_hierarchical(radius_cutoff=0.25, member_cutoff=50)

    t = _timed(run, args.steps, args.warmup)
    g = engine.build_graph(Xd, 0.0625, "f32_band")
    row_ptr, col = g.to_csr()
    g.release()
    nnz = int(col.shape[0])
    engine.edge_similarity(row_ptr, col)
    t0 = time.perf_counter()
    for _ in range(3):
        engine.edge_similarity(row_ptr, col)  # synchronous: returns after the kernel
    t_w = (time.perf_counter() - t0) / 3 * 1e3
    t0 = time.perf_counter()
    engine.similarity_mst(row_ptr, col)
    t_mst = time.perf_counter() - t0
    log(f"hierarchy_mst: {t * 1e3:.1f} ms per hierarchy; weights {t_w:.2f} ms for {nnz} entries; "
        f"weights + Prim {t_mst * 1e3:.1f} ms")
    cpu = None
    if not args.no_cpu:
        # the reference's procedure (set-intersection weights, heapq Prim, merge
        # hierarchy) as restated by the oracle, on the 3000 points nearest to a
        # seeded point of the same sample: the same density, so the same list
        # lengths; its cost grows about linearly in n at fixed density
        from oracle import cnn_oracle as O

        centre = Xs[np.random.default_rng(2).integers(0, m)]
        near = np.argsort(((Xs.astype(np.float64) - centre) ** 2).sum(1))[:3000]
        S = np.ascontiguousarray(Xs[np.sort(near)], dtype=np.float64)
        t0 = time.perf_counter()
        neigh = O.neighbourhoods(S, 0.25)
        O.mst_hierarchy(len(S), O.mst_prim(neigh), 50)
        ts = time.perf_counter() - t0
        deg = float(np.mean([len(x) for x in neigh]))
        cpu = {"value": len(S) / ts, "unit": "points/s", "cores": 1, "kind": "port",
               "sample": f"3000 nearest points of a seeded centre (mean list length {deg:.0f}, the full "
                         f"sample's {nnz / m:.0f}), oracle restatement of the reference's weights, heapq Prim "
                         f"and merge hierarchy in {ts:.1f} s"}
        log(f"cpu sample: {cpu['sample']}")
    return _line(m / t, t * 1e3, args.steps, args.warmup,
                 {"workload": "hierarchy_mst", "n": m, "d": 3, "radius_cutoff": 0.25, "member_cutoff": 50},
                 {"nnz": nnz, "weights_kernel_ms": t_w, "weights_entries_per_s": nnz / (t_w * 1e-3),
                  "weights_plus_prim_ms": t_mst * 1e3, "tree_edges": len(cl.hierarchical_fitter.last_tree[0]),
                  "cpu_baseline": cpu})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="md_basins_3d")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 fits and the Clustering.fit timing (profiling runs)")
    ap.add_argument("--requests", type=int, default=300, help="serving workload: requests per arrival rate")
    ap.add_argument("--served-requests", type=int, default=300,
                    help="requests per arrival rate of the served block on the default line")
    ap.add_argument("--no-extra", action="store_true",
                    help="default workload: leave out the features_64d and served blocks (profiling runs)")
    ap.add_argument("--slots", type=int, default=4, help="serving workload: concurrent slots")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run a plain device-to-device copy of MIB MiB three times and exit; its rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE readings calibrate the traffic counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: every box in this setup has a single MI355X "
              f"(got --gpus {args.gpus}); run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe:
        nbytes = args.copy_probe << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 3 copies of {nbytes} bytes")
        return

    if args.workload == "serving":
        print(json.dumps(serving_bench(args)))
        return
    if args.workload == "scan_2d":
        print(json.dumps(scan_bench(args)))
        return
    if args.workload == "hierarchy_repeat":
        print(json.dumps(hierarchy_repeat_bench(args)))
        return
    if args.workload == "hierarchy_mst":
        print(json.dumps(hierarchy_mst_bench(args)))
        return
    if args.workload == "many_small":
        print(json.dumps(batch_bench(args)))
        return

    X, p = workload(args.workload)
    n, d = X.shape
    Xd = torch.from_numpy(X).cuda()
    log(f"workload {args.workload}: n={n} d={d} r={p['radius_cutoff']:.6g} c={p['cnn_cutoff']}")
    t_band, lab_band, gs_band, ls_band = run_fits(Xd, p, "f32_band", args.steps, args.warmup)
    log(f"f32_band: {t_band * 1e3:.2f} ms/fit")
    out_plain, identical, ph_plain = None, None, None
    if not args.no_plain:
        t_plain, lab_plain, gs_plain, ls_plain = run_fits(Xd, p, "f64", args.steps, max(1, args.warmup // 2))
        log(f"plain f64: {t_plain * 1e3:.2f} ms/fit")
        identical = bool(np.array_equal(lab_band, lab_plain))
        out_plain = {"points_per_s": n / t_plain, "ms_per_step": t_plain * 1e3,
                     "distance_roofline": distance_roofline(gs_plain, d, "f64", args.workload)}
        ph_plain = phases(gs_plain, ls_plain)

    # host-buffer form: X starts in (pinned) host memory, the copy to HBM is timed with the fit
    import torch
    from commonnnclustering_amd import autotune, engine

    Xh = torch.from_numpy(X).pin_memory()
    r2 = float(p["radius_cutoff"]) ** 2
    c = int(p["cnn_cutoff"])

    def h2d_fit():
        engine.fit(Xh.cuda(non_blocking=True), r2, c + 2, c + 1, precision="f32_band",
                   member_cutoff=p["member_cutoff"])

    t_h2d = _timed(h2d_fit, max(1, args.steps // 2), 1)
    log(f"f32_band from host memory (H2D included): {t_h2d * 1e3:.2f} ms/fit")

    # the drop-in API: Clustering(points).fit(r, c, member_cutoff) with host input
    # (coordinates cached on the device after the first fit; labels copied back,
    # label meta data and the Record built on the host)
    from commonnnclustering_amd import Clustering

    api = {}
    for recipe in (() if args.no_plain else ("coordinates", "coordinates_f32")):
        cl = Clustering(X, recipe=recipe)

        def api_fit():
            cl.fit(p["radius_cutoff"], p["cnn_cutoff"], member_cutoff=p["member_cutoff"], v=False)

        api[recipe] = _timed(api_fit, max(1, args.steps // 2), 1) * 1e3
        if not np.array_equal(cl.labels, lab_band):
            raise SystemExit(f"Clustering.fit ({recipe}) labels differ from the engine's")
    if api:
        log("Clustering.fit: " + ", ".join(f"{k} {v:.2f} ms" for k, v in api.items()))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(X, p, args.cpu_seconds)
        log(f"cpu baseline: {cpu['sample']}")

    # the roofline line describes the longest single kernel of the fit: the
    # distance kernel, the CSR emit (when the fit emits lists) or the label pass's
    # full round
    dist_rf = distance_roofline(gs_band, d, "f32_band", args.workload)
    emit_rf = emit_roofline(gs_band, n, args.workload) if mean(gs_band, "ms_emit") > 0 else None
    su_rf = su_rooflines(ls_band, args.workload)
    dominant = dist_rf
    if emit_rf and emit_rf["ms_per_launch"] > dominant["ms_per_launch"]:
        dominant = emit_rf
    rb = su_rf["round"]
    if rb["ms_per_launch"] > dominant["ms_per_launch"] and rb["achieved"] is not None:
        dominant = rb
    rows_bad, entries = neighbour_mismatches(Xd, r2)
    log(f"neighbour rows differing between f32_band and plain f64: {rows_bad} of {n}")
    extra = {}
    if not args.no_extra:
        extra["features_64d"] = features_block(args)
        extra["served"] = served_block(args, count=args.served_requests)

    out = {
        "metric": METRIC,
        "value": n / t_band,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_band * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded numpy generator)",
        "config": {"workload": args.workload, "n": int(n), "d": int(d), "dtype_in": str(X.dtype),
                   "radius_cutoff": float(p["radius_cutoff"]), "cnn_cutoff": int(p["cnn_cutoff"]),
                   "member_cutoff": int(p["member_cutoff"])},
        "plain_f64": out_plain,
        "band_vs_plain": (out_plain["ms_per_step"] / (t_band * 1e3)) if out_plain else None,
        "labels_identical": identical,
        "n_clusters": int(ls_band[-1]["n_clusters"]),
        "nnz": int(gs_band[-1]["nnz"]),
        "band_hits": int(gs_band[-1]["band_hits"]),
        "band_flips": int(gs_band[-1]["band_flips"]),
        "max_fp32_abs_err_band": float(gs_band[-1]["max_abs_err_band"]),
        "with_h2d": {"points_per_s": n / t_h2d, "ms_per_step": t_h2d * 1e3,
                     "note": "X copied from pinned host memory inside the timed region; not value"},
        "clustering_fit_ms": api,
        "phase_ms": phases(gs_band, ls_band),
        "phase_ms_plain": ph_plain,
        "label_stats": ls_band[-1],
        "tuning": {"f32_band": autotune.current(d, n, "f32_band"), "f64": autotune.current(d, n, "f64"),
                   "source": "commonnnclustering_amd/tuning_cache.json (python -m commonnnclustering_amd.autotune)"},
        "neighbour_rows_differing_from_plain_f64": rows_bad,
        "neighbour_entries": entries,
        "roofline": dominant,
        "kernel_rooflines": {"distance": dist_rf, "emit": emit_rf},
        "similarity_rooflines": su_rf,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
