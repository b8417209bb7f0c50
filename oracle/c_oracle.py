"""ctypes wrapper of the C restatement (``cnn_oracle.c``); test infrastructure only."""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64


def build(force: bool = False) -> str:
    """Compile ``liboracle.so`` with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_neighbours_count.argtypes = [_f64p, _i64, ctypes.c_int, ctypes.c_double, ctypes.c_int, _i64p]
        L.oracle_neighbours_fill.argtypes = [_f64p, _i64, ctypes.c_int, ctypes.c_double, ctypes.c_int, _i64p, _i32p]
        L.oracle_labels_bfs.argtypes = [_i64p, _i32p, _i64, _i64, _i64, _u8p, _i64, _i64p]
        L.oracle_labels_cc.argtypes = [_i64p, _i32p, _i64, _i64, _i64, _u8p, _i64, _i64p]
        L.oracle_merge_work.argtypes = [_i64p, _i32p, _i64, _i64, _i64, _u8p,
                                        ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)]
        L.oracle_sort_by_size.argtypes = [_i64p, _i64, _i64, _i64, ctypes.POINTER(_i64)]
        L.oracle_check_similarity.argtypes = [_i32p, _i64, _i32p, _i64, _i64]
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_set_num_threads.argtypes = [ctypes.c_int]
        L.oracle_neighbours_bf.argtypes = [_f64p, _i64, ctypes.c_int, ctypes.c_double, ctypes.c_void_p, _i64,
                                           _i64p, ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_free.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def num_threads() -> int:
    return lib().oracle_num_threads()


def set_num_threads(k: int) -> None:
    """Threads of the OpenMP loops from now on (all host threads by default)."""
    lib().oracle_set_num_threads(int(k))


def neighbours_csr(X, radius_cutoff: float, use_grid: bool = True):
    """(row_ptr int64[n+1], col int32[nnz]) of N(i) = {j : s_ij <= r*r}, rows ascending."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, d = X.shape
    r = float(radius_cutoff)
    r2 = r * r
    deg = np.zeros(n, dtype=np.int64)
    rc = lib().oracle_neighbours_count(X, n, d, r2, int(use_grid), deg)
    if rc != 0:
        raise MemoryError("oracle_neighbours_count failed")
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(deg, out=row_ptr[1:])
    col = np.empty(int(row_ptr[-1]), dtype=np.int32)
    rc = lib().oracle_neighbours_fill(X, n, d, r2, int(use_grid), row_ptr, col)
    if rc != 0:
        raise MemoryError("oracle_neighbours_fill failed")
    return row_ptr, col


def labels_from_csr(row_ptr, col, c_sim: int, m_core: int, consider=None,
                    current_start: int = 1, labels=None, method: str = "cc"):
    n = len(row_ptr) - 1
    if consider is None:
        consider = np.ones(n, dtype=np.uint8)
    consider = np.ascontiguousarray(consider, dtype=np.uint8).copy()
    if labels is None:
        labels = np.zeros(n, dtype=np.int64)
    fn = lib().oracle_labels_cc if method == "cc" else lib().oracle_labels_bfs
    rc = fn(np.ascontiguousarray(row_ptr, dtype=np.int64), np.ascontiguousarray(col, dtype=np.int32),
            n, int(c_sim), int(m_core), consider, int(current_start), labels)
    if rc != 0:
        raise MemoryError("oracle labelling failed")
    return labels, consider


def merge_work(row_ptr, col, c_sim: int, m_core: int, consider=None):
    n = len(row_ptr) - 1
    if consider is None:
        consider = np.ones(n, dtype=np.uint8)
    el, cand, pas = _i64(), _i64(), _i64()
    lib().oracle_merge_work(np.ascontiguousarray(row_ptr, dtype=np.int64),
                            np.ascontiguousarray(col, dtype=np.int32), n, int(c_sim), int(m_core),
                            np.ascontiguousarray(consider, dtype=np.uint8),
                            ctypes.byref(el), ctypes.byref(cand), ctypes.byref(pas))
    return el.value, cand.value, pas.value


def sort_by_size(labels, member_cutoff=None, max_clusters=None):
    """In place on an int64 array; returns the number of surviving clusters."""
    if member_cutoff is None:
        member_cutoff = 2
    if max_clusters is None:
        max_clusters = -1
    k = _i64()
    lib().oracle_sort_by_size(labels, len(labels), int(member_cutoff), int(max_clusters), ctypes.byref(k))
    return k.value


def fit(X, radius_cutoff, cnn_cutoff, member_cutoff=None, max_clusters=None,
        use_grid=True, method="cc", sort=True):
    """End-to-end fit (cluster.pyx:256-392 restated): labels int64[n]."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    row_ptr, col = neighbours_csr(X, radius_cutoff, use_grid)
    c = int(cnn_cutoff)
    labels, _ = labels_from_csr(row_ptr, col, c + 2, c + 1, method=method)
    if sort:
        sort_by_size(labels, member_cutoff, max_clusters)
    return labels


def neighbours_bf(X, radius_cutoff: float = None, rows=None, r2: float = None):
    """Brute-force N(i) (rows ascending) for every point, or for the listed rows:
    (row_ptr int64[len(rows)+1], col int32[nnz]).  Exact early exit (see
    oracle_neighbours_bf); the squared radius is r*r unless r2 is given."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, d = X.shape
    if r2 is None:
        r = float(radius_cutoff)
        r2 = r * r
    L = lib()
    if rows is None:
        rows_arr, n_rows, rows_p = None, n, None
    else:
        rows_arr = np.ascontiguousarray(rows, dtype=np.int64)
        if rows_arr.size and (rows_arr.min() < 0 or rows_arr.max() >= n):
            raise IndexError("row index out of range")
        n_rows = rows_arr.size
        rows_p = rows_arr.ctypes.data_as(ctypes.c_void_p)
    row_ptr = np.zeros(n_rows + 1, dtype=np.int64)
    out = ctypes.c_void_p()
    rc = L.oracle_neighbours_bf(X, n, d, float(r2), rows_p, n_rows, row_ptr, ctypes.byref(out))
    if rc != 0:
        raise MemoryError("oracle_neighbours_bf failed")
    nnz = int(row_ptr[-1])
    try:
        col = np.empty(nnz, dtype=np.int32)
        if nnz:
            ctypes.memmove(col.ctypes.data, out.value, nnz * 4)
    finally:
        L.oracle_free(out)
    return row_ptr, col
