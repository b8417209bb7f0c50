"""ctypes binding of libcnn_hip.so (the C-ABI declared in include/cnn_gpu.h).

The product path has exactly one implementation: the HIP kernels behind this
library.  Loading fails loudly when the library is missing; there is no CPU
fallback.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcnn_hip.so")

CNN_OK = 0
CNN_ERR_INVALID = -1
CNN_ERR_OOM = -2
CNN_ERR_HIP = -3
CNN_ERR_CAPACITY = -4
CNN_ERR_UNSUPPORTED = -5

CNN_F32 = 0
CNN_F64 = 1

CNN_SIM_FORM_MERGE = 1
CNN_SIM_FORM_TILE_WORDS = 2
CNN_SIM_FORM_ROW_WORDS = 3

CNN_PLAIN_F64 = 0
CNN_F32_BAND = 1

CNN_SIM_AUTO = 0
CNN_SIM_MERGE = 1
CNN_SIM_BITSET = 2

PRECISIONS = {"f64": CNN_PLAIN_F64, "plain_f64": CNN_PLAIN_F64, "plain": CNN_PLAIN_F64,
              "f32_band": CNN_F32_BAND, "f32": CNN_F32_BAND, "band": CNN_F32_BAND}

# Every symbol include/cnn_gpu.h declares.
EXPORTED = [
    "cnn_ctx_create", "cnn_ctx_destroy", "cnn_last_error", "cnn_ctx_set_stream",
    "cnn_ctx_get_stream", "cnn_ctx_synchronize", "cnn_ctx_workspace_bytes", "cnn_ctx_trim",
    "cnn_graph_build", "cnn_graph_destroy", "cnn_graph_nnz", "cnn_graph_stats_get",
    "cnn_graph_bytes", "cnn_graph_export_csr", "cnn_radius_csr", "cnn_graph_rethreshold",
    "cnn_graph_labels", "cnn_labels_from_csr", "cnn_sort_by_size", "cnn_fit", "cnn_graph_fit",
    "cnn_fit_batch", "cnn_fit_scan", "cnn_predict", "cnn_precomputed_csr", "cnn_graph_build_periodic",
    "cnn_csr_edge_similarity", "cnn_similarity_mst", "cnn_version", "cnn_labels_from_csr_noalloc",
    "cnn_tune_set", "cnn_tune_get", "cnn_tune_clear", "cnn_graph_scan_create", "cnn_scan_launch",
    "cnn_scan_destroy",
]


class CNNError(RuntimeError):
    """A C-ABI call returned a non-zero cnn_status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[cnn_status {code}] {message}")
        self.code = code


class CNNCapacityError(CNNError):
    pass


class GraphStats(ctypes.Structure):
    _fields_ = [
        ("pairs_evaluated", ctypes.c_int64),
        ("band_hits", ctypes.c_int64),
        ("band_flips", ctypes.c_int64),
        ("max_abs_err_band", ctypes.c_double),
        ("tile_pairs", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("ms_prep", ctypes.c_float),
        ("ms_distance", ctypes.c_float),
        ("ms_emit", ctypes.c_float),
        ("distance_launches", ctypes.c_int32),
        ("layout", ctypes.c_int32),
        ("tile_pairs_computed", ctypes.c_int64),
        ("ms_distance_kernel", ctypes.c_float),
    ]

    def to_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class TuneParams(ctypes.Structure):
    """cnn_tune_params: launch shapes and switches of one problem class."""

    _fields_ = [
        ("mfma_col_chunk", ctypes.c_int32),
        ("wide_col_chunk", ctypes.c_int32),
        ("emit_waves", ctypes.c_int32),
        ("root_passes", ctypes.c_int32),
        ("words_min_fill", ctypes.c_int32),
    ]

    def to_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class LabelStats(ctypes.Structure):
    _fields_ = [
        ("n_core", ctypes.c_int64),
        ("n_components", ctypes.c_int64),
        ("n_clusters", ctypes.c_int64),
        ("tests", ctypes.c_int64),
        ("tests_passed", ctypes.c_int64),
        ("sampled_components", ctypes.c_int64),
        ("rows_scanned", ctypes.c_int64),
        ("ms_similarity", ctypes.c_float),
        ("ms_labels", ctypes.c_float),
        ("sim_form", ctypes.c_int32),
        ("ms_hook", ctypes.c_float),
        ("ms_round", ctypes.c_float),
    ]

    def to_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_dbl = ctypes.c_double
_i64p = ctypes.POINTER(ctypes.c_int64)
_lib = None
_lock = threading.Lock()


def _declare(L):
    c = ctypes.c_int
    sig = {
        "cnn_ctx_create": (c, [c, ctypes.POINTER(_vp)]),
        "cnn_ctx_destroy": (c, [_vp]),
        "cnn_last_error": (ctypes.c_char_p, [_vp]),
        "cnn_ctx_set_stream": (c, [_vp, _vp]),
        "cnn_ctx_get_stream": (_vp, [_vp]),
        "cnn_ctx_synchronize": (c, [_vp]),
        "cnn_ctx_workspace_bytes": (c, [_vp, _i64p]),
        "cnn_ctx_trim": (c, [_vp]),
        "cnn_graph_build": (c, [_vp, _vp, _i64, _i32, _i32, _dbl, _i32, ctypes.POINTER(_vp)]),
        "cnn_graph_destroy": (c, [_vp]),
        "cnn_graph_nnz": (c, [_vp, _i64p]),
        "cnn_graph_stats_get": (c, [_vp, ctypes.POINTER(GraphStats)]),
        "cnn_graph_bytes": (c, [_vp, _i64p]),
        "cnn_graph_export_csr": (c, [_vp, _vp, _vp, _vp, _i64]),
        "cnn_radius_csr": (c, [_vp, _vp, _i64, _i32, _i32, _dbl, _i32, _vp, _vp, _i64, _i64p,
                               ctypes.POINTER(GraphStats)]),
        "cnn_graph_rethreshold": (c, [_vp, _vp, _dbl, _i32, ctypes.POINTER(_vp)]),
        "cnn_graph_labels": (c, [_vp, _vp, _i32, _i32, _vp, _i64, _i32, _vp, ctypes.POINTER(LabelStats)]),
        "cnn_labels_from_csr": (c, [_vp, _vp, _vp, _i64, _i32, _i32, _vp, _i64, _i32, _vp,
                                    ctypes.POINTER(LabelStats)]),
        "cnn_labels_from_csr_noalloc": (c, [_vp, _vp, _vp, _i64, _i32, _i32, _vp, _i64, _i32, _vp,
                                            ctypes.POINTER(LabelStats)]),
        "cnn_sort_by_size": (c, [_vp, _vp, _i64, _i64, _i64, _i64p]),
        "cnn_tune_set": (c, [_i32, _i32, _i64, _i64, _i32, ctypes.POINTER(TuneParams)]),
        "cnn_tune_get": (c, [_i32, _i64, _i32, ctypes.POINTER(TuneParams)]),
        "cnn_tune_clear": (c, []),
        "cnn_graph_scan_create": (c, [_vp, _vp, _vp, _i32, _vp, _vp, _i32, _i64, _i64, _vp, ctypes.POINTER(_vp)]),
        "cnn_scan_launch": (c, [_vp, _vp, _vp]),
        "cnn_scan_destroy": (c, [_vp]),
        "cnn_fit": (c, [_vp, _vp, _i64, _i32, _i32, _dbl, _i32, _i32, _i32, _i32, _i64, _i64, _vp,
                        ctypes.POINTER(_vp), ctypes.POINTER(GraphStats), ctypes.POINTER(LabelStats)]),
        "cnn_graph_fit": (c, [_vp, _vp, _i32, _i32, _i32, _i64, _i64, _vp, ctypes.POINTER(LabelStats)]),
        "cnn_fit_batch": (c, [_vp, _vp, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _i32, _i64, _i64, _vp,
                              ctypes.POINTER(GraphStats), ctypes.POINTER(LabelStats)]),
        "cnn_fit_scan": (c, [_vp, _vp, _i64, _i32, _i32, _vp, _i32, _vp, _vp, _i32, _i32, _i64, _i64, _vp,
                             ctypes.POINTER(GraphStats), _vp]),
        "cnn_predict": (c, [_vp, _vp, _i64, _vp, _i64, _i32, _i32, ctypes.c_double, _i32, _i32, _i32, _vp, _vp,
                            _i64, _vp, _vp, ctypes.POINTER(_i64)]),
        "cnn_precomputed_csr": (c, [_vp, _vp, _i64, _i32, ctypes.c_double, _vp, _vp, _i64,
                                    ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
        "cnn_graph_build_periodic": (c, [_vp, _vp, _i64, _i32, _i32, ctypes.c_double, _vp, ctypes.POINTER(_vp)]),
        "cnn_csr_edge_similarity": (c, [_vp, _vp, _vp, _i64, _vp]),
        "cnn_similarity_mst": (c, [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _i64, ctypes.POINTER(_i64)]),
        "cnn_version": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def lib():
    """Load libcnn_hip.so; raises ImportError when it has not been built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(
                        f"{LIB_PATH} is missing: build it with `python -m commonnnclustering_amd._build` "
                        "(the engine has no CPU fallback)")
                L = ctypes.CDLL(LIB_PATH)
                _declare(L)
                _lib = L
                # tuned launch shapes (commonnnclustering_amd/tuning_cache.json), when present
                from . import autotune

                autotune.load_cache()
    return _lib


class Context:
    """One cnn_ctx (device, stream, workspace arena)."""

    def __init__(self, device: int = 0):
        self.device = int(device)
        h = _vp()
        rc = lib().cnn_ctx_create(self.device, ctypes.byref(h))
        if rc != CNN_OK:
            raise CNNError(rc, f"cnn_ctx_create(device={device}) failed")
        self.handle = h

    def check(self, rc: int, what: str = "") -> None:
        if rc == CNN_OK:
            return
        msg = lib().cnn_last_error(self.handle)
        msg = msg.decode() if msg else ""
        cls = CNNCapacityError if rc == CNN_ERR_CAPACITY else CNNError
        raise cls(rc, f"{what}: {msg}" if what else msg)

    def use_stream(self, stream) -> None:
        """Enqueue on a torch.cuda.Stream / raw hipStream_t (None or 0: default stream)."""
        if stream is None:
            ptr = None
        elif isinstance(stream, int):
            ptr = stream
        else:
            ptr = stream.cuda_stream
        self.check(lib().cnn_ctx_set_stream(self.handle, ptr), "cnn_ctx_set_stream")

    def synchronize(self) -> None:
        self.check(lib().cnn_ctx_synchronize(self.handle), "cnn_ctx_synchronize")

    def workspace_bytes(self) -> int:
        v = _i64()
        self.check(lib().cnn_ctx_workspace_bytes(self.handle, ctypes.byref(v)))
        return v.value

    def trim(self) -> None:
        self.check(lib().cnn_ctx_trim(self.handle), "cnn_ctx_trim")

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().cnn_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts: dict = {}


def get_context(device=None) -> Context:
    """Per-thread, per-device cached context bound to torch's current stream."""
    import torch

    if device is None:
        device = torch.cuda.current_device()
    if isinstance(device, torch.device):
        device = device.index if device.index is not None else torch.cuda.current_device()
    key = (threading.get_ident(), int(device))
    ctx = _contexts.get(key)
    if ctx is None:
        ctx = Context(int(device))
        _contexts[key] = ctx
    ctx.use_stream(torch.cuda.current_stream(int(device)))
    return ctx


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def dtype_code(t) -> int:
    import torch

    if t.dtype == torch.float32:
        return CNN_F32
    if t.dtype == torch.float64:
        return CNN_F64
    raise TypeError(f"coordinates must be float32 or float64, got {t.dtype}")


def precision_code(p) -> int:
    if isinstance(p, (int, np.integer)):
        if int(p) not in (CNN_PLAIN_F64, CNN_F32_BAND):
            raise ValueError(f"unknown precision {p}")
        return int(p)
    try:
        return PRECISIONS[str(p).lower()]
    except KeyError:
        raise ValueError(f"unknown precision {p!r}; use one of {sorted(PRECISIONS)}") from None
