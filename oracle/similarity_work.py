"""Algorithmic work of the similarity and union-find phases (test
infrastructure: measurement context for the bench line).

For a workload's reference neighbour lists (caller order), the canonical
two-pointer merge with early exit at c + 2 (_types.pyx:2654-2706) over every
candidate edge (i < j, both core) consumes l_i + l_j elements; the similarity
phase's algorithmic bytes are B_S = 4 * sum(l_i + l_j), and the union-find's
B_U = 8 |E_pass| (endpoint reads) + 8 n (parents) + 8 n (labels) + 16 n
(sort_by_size keys).  Work a union-find skips is not subtracted.

    python oracle/similarity_work.py md_basins_3d features_64d

writes profiles/similarity_work_<workload>.json.
"""

from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)


def work(name: str) -> dict:
    from commonnnclustering_amd import synthetic
    from oracle import c_oracle as C

    X, p = getattr(synthetic, name)()
    n, d = X.shape
    r, c = p["radius_cutoff"], p["cnn_cutoff"]
    t0 = time.perf_counter()
    rp, col = C.neighbours_csr(X, r) if d <= 3 else C.neighbours_bf(X, r)
    elements, candidates, passing = C.merge_work(rp, col, c + 2, c + 1)
    return {"workload": name, "n": int(n), "d": int(d), "radius_cutoff": float(r), "cnn_cutoff": int(c),
            "nnz": int(rp[-1]), "candidate_edges": int(candidates), "passing_edges": int(passing),
            "merge_elements": int(elements), "B_S_bytes": 4 * int(elements),
            "B_U_bytes": 8 * int(passing) + 32 * int(n),
            "seconds": round(time.perf_counter() - t0, 1),
            "command": "python oracle/similarity_work.py " + name}


def main(names):
    for name in names:
        out = work(name)
        path = os.path.join(ROOT, "profiles", f"similarity_work_{name}.json")
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps(out))


if __name__ == "__main__":
    main(sys.argv[1:] or ["md_basins_3d", "features_64d"])
