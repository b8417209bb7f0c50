"""The C-ABI library loads and exports every function include/cnn_gpu.h
declares, and the Python binding declares the same set (no GPU calls)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cnn_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cnn_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "cnn_fit" in names and "cnn_labels_from_csr" in names and "cnn_fit_scan" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from commonnnclustering_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libcnn_hip.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    from commonnnclustering_amd import _lib

    assert sorted(_lib.EXPORTED) == declared_functions()


def test_version_string():
    from commonnnclustering_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libcnn_hip.so not built")
    v = _lib.lib().cnn_version()
    assert v and "gfx950" in v.decode()
