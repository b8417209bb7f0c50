"""Extract the reference's known-answer vectors into tests/golden/kats.json.

Reads the reference's documentation and test files as TEXT (nothing is
imported or executed from the reference).  The fixture holds data only:
inputs, parameters and expected outputs, each with its source file:line.

Also stores the backbone-dihedral data file of the periodic-metric test as
tests/golden/backbone_dihedrals.npy (loaded with allow_pickle=False).

Usage: python tests/golden/make_golden.py [/path/to/reference]
"""

import ast
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kats.json")


def read(rel):
    with open(os.path.join(REF, rel)) as fh:
        return fh.read()


def line_of(text, needle):
    return text[: text.index(needle)].count("\n") + 1


def quickstart():
    text = read("README.md")
    block = text[text.index("data_points = ["):]
    block = block[: block.index("]\n") + 1]
    pts = ast.literal_eval(re.sub(r"#.*", "", block.split("=", 1)[1]).replace("...", ""))
    m = re.search(r"fit\(radius_cutoff=([\d.]+), cnn_cutoff=(\d+)", text)
    lab = re.search(r"array\((\[[\d, ]+\])\)", text)
    return {
        "source": f"README.md:{line_of(text, 'data_points = [')}-{line_of(text, lab.group(0))}",
        "points": pts,
        "radius_cutoff": float(m.group(1)),
        "cnn_cutoff": int(m.group(2)),
        "labels": ast.literal_eval(lab.group(1)),
    }


def basic_usage():
    nb = json.loads(read("docsrc/tutorial/basic_usage.ipynb"))
    raw = read("docsrc/tutorial/basic_usage.ipynb")
    out = {}
    for cell in nb["cells"]:
        if cell["cell_type"] != "code":
            continue
        src = "".join(cell["source"])
        if src.strip() == "clustering.labels":
            for o in cell.get("outputs", []):
                txt = "".join(o.get("data", {}).get("text/plain", []))
                if txt.startswith("array"):
                    out["labels_r2"] = ast.literal_eval(txt[len("array("):-1])
        if "fit(radius_cutoff=2.0" in src or "fit(radius_cutoff=1.5" in src:
            for o in cell.get("outputs", []):
                for line in "".join(o.get("text", [])).splitlines():
                    if line.startswith("12 "):
                        f = line.split()
                        out.setdefault("records", []).append({
                            "n_points": int(f[0]), "radius_cutoff": float(f[1]), "cnn_cutoff": int(f[2]),
                            "n_clusters": int(f[5]), "ratio_largest": float(f[6]), "ratio_noise": float(f[7])})
    seen = []
    for r in out.get("records", []):
        if r not in seen:
            seen.append(r)
    out["records"] = seen
    out["source"] = f"docsrc/tutorial/basic_usage.ipynb:{line_of(raw, 'array([1, 1, 1, 1, 1, 1, 1, 0, 0, 2, 2, 2])')}"
    return out


def sklearn_records():
    raw = read("docsrc/tutorial/scikit_learn_datasets.ipynb")
    nb = json.loads(raw)
    recs = []
    for cell in nb["cells"]:
        src = "".join(cell.get("source", []))
        if cell["cell_type"] == "code" and "clustering = cluster.Clustering(data)" in src and \
                "clustering.fit(**params)" in src:
            for o in cell.get("outputs", []):
                for line in "".join(o.get("text", [])).splitlines():
                    if line.startswith("2000 "):
                        f = line.split()
                        recs.append({"n_points": 2000, "radius_cutoff": float(f[1]), "cnn_cutoff": int(f[2]),
                                     "member_cutoff": int(f[3]), "n_clusters": int(f[5]),
                                     "ratio_largest": float(f[6]), "ratio_noise": float(f[7]),
                                     "execution_time_s": f[8]})
            break
    names = ["circles", "moons", "varied", "aniso", "blobs", "None"]
    for r, name in zip(recs, names):
        r["dataset"] = name
    return {"source": "docsrc/tutorial/scikit_learn_datasets.ipynb:384-414 (data generation :205-270, "
                      "parameters :285-322)", "records": recs}


def similarity_cases():
    text = read("test/unit/test_check_similarity.py")
    block = text[text.index('"members_a,members_b,c,is_sorted,expected"'):]
    block = block[block.index("["): block.index("]\n    )") + 1]
    cases = ast.literal_eval(block)
    return {"source": f"test/unit/test_check_similarity.py:{line_of(text, 'members_a,members_b,c,is_sorted,expected')}",
            "cases": [{"a": a, "b": b, "c": c, "sorted": s, "expected": e} for a, b, c, s, e in cases]}


def sort_cases():
    text = read("test/unit/test_types.py")
    i = text.index("def test_sort_by_size")
    head = text[:i]
    j = head.rindex('"kwargs,expected"')
    block = head[head.index("[", j): head.rindex("]") + 1]
    cases = ast.literal_eval(block)
    labels = ast.literal_eval(re.search(r"np\.array\((\[[\d, ]+\]), dtype=P_AINDEX", text[i:]).group(1))
    return {"source": f"test/unit/test_types.py:{line_of(text, chr(34) + 'kwargs,expected' + chr(34))}-"
                      f"{line_of(text, 'def test_sort_by_size') + 10}",
            "labels": labels, "cases": [{"kwargs": k, "expected": e} for k, e in cases]}


def params_repr():
    text = read("test/unit/test_types/test_create_params_0_1_1_.txt")
    return {"source": "test/unit/test_types/test_create_params_0_1_1_.txt:1",
            "args": [0.1, 1, 1], "repr": text.strip().splitlines()[0]}


def metric_points():
    text = read("test/unit/test_metric.py")
    i = text.index("InputDataExtComponentsMemoryview,\n                np.array([")
    blk = text[i:]
    arrs = re.findall(r"np\.array\(\[(.*?)\], order", blk, re.S)[:2]
    data = [ast.literal_eval("[" + a + "]") for a in arrs]
    return {"source": f"test/unit/test_metric.py:{text[:i].count(chr(10)) + 1}",
            "data": data[0], "other": data[1], "significant": 12}


def _literal(node):
    """A literal node, or the first argument of an np.array(...) call, as data."""
    if isinstance(node, ast.Call):
        return ast.literal_eval(node.args[0])
    return ast.literal_eval(node)


def _parametrize_cases(tree, func_name):
    for node in ast.walk(tree):
        if isinstance(node, ast.FunctionDef) and node.name == func_name:
            for dec in node.decorator_list:
                if isinstance(dec, ast.Call) and getattr(dec.func, "attr", "") == "parametrize":
                    return dec.args[1].elts, node.lineno
    raise KeyError(func_name)


def isolate_cases():
    """test_isolate (test/unit/test_clustering.py) with its regression reports."""
    rel = "test/unit/test_clustering.py"
    tree = ast.parse(read(rel))
    cases, lineno = _parametrize_cases(tree, "test_isolate")
    reports = sorted(os.listdir(os.path.join(REF, "test/unit/test_clustering")))
    out = []
    for k, case in enumerate(cases):
        _, data, meta, labels, root, parent = case.elts
        name = next(r for r in reports if f"_data{k}_" in r)
        out.append({"data": _literal(data), "meta": _literal(meta), "labels": _literal(labels),
                    "root_indices": _literal(root), "parent_indices": _literal(parent),
                    "report": read(f"test/unit/test_clustering/{name}"),
                    "report_source": f"test/unit/test_clustering/{name}"})
    return {"source": f"{rel}:{lineno}", "cases": out}


def reel_cases():
    """test_reel (test/unit/test_clustering.py) and its fixture hierarchy
    (test/conftest.py make_hierarchical_clustering_a)."""
    rel = "test/unit/test_clustering.py"
    cases, lineno = _parametrize_cases(ast.parse(read(rel)), "test_reel")
    conf = read("test/conftest.py")
    fixture = next(n for n in ast.walk(ast.parse(conf))
                   if isinstance(n, ast.FunctionDef) and n.name == "make_hierarchical_clustering_a")
    calls = [n for n in ast.walk(fixture) if isinstance(n, ast.Call) and getattr(n.func, "attr", "") == "array"]
    arrays = [_literal(n) for n in sorted(calls, key=lambda n: (n.lineno, n.col_offset))]
    # fixture order: root labels; child 1 labels, root, parent; child 1.2 labels, root, parent;
    # children 0, 1, 2 of the root and of child 1 are added empty
    root_labels, c1_labels, c1_root, c1_parent, c12_labels, c12_root, c12_parent = arrays
    return {"source": f"{rel}:{lineno}; test/conftest.py:{fixture.lineno}",
            "root_labels": root_labels,
            "empty_children": [[0], [1], [2], [1, 0], [1, 1], [1, 2]],
            "nodes": [{"path": [1], "labels": c1_labels, "root": c1_root, "parent": c1_parent},
                      {"path": [1, 2], "labels": c12_labels, "root": c12_root, "parent": c12_parent}],
            "cases": [{"depth": d, "expected": e} for _, d, e in (ast.literal_eval(c) for c in cases)]}


def main():
    kats = {
        "quickstart": quickstart(),
        "basic_usage": basic_usage(),
        "sklearn_records": sklearn_records(),
        "similarity": similarity_cases(),
        "sort_by_size": sort_cases(),
        "params_repr": params_repr(),
        "metric": metric_points(),
        "isolate": isolate_cases(),
        "reel": reel_cases(),
    }
    with open(OUT, "w") as fh:
        json.dump(kats, fh, indent=1)
    print(f"wrote {OUT}")
    # data file of the reference's periodic-metric test (test_procedures.py:175-215),
    # loaded without unpickling and stored as a fixture
    import numpy as np

    src = os.path.join(REF, "test/integration/test_procedures/backbone_dihedrals.npy")
    dst = os.path.join(os.path.dirname(OUT), "backbone_dihedrals.npy")
    np.save(dst, np.ascontiguousarray(np.load(src, allow_pickle=False)))
    print(f"wrote {dst}")


if __name__ == "__main__":
    main()
