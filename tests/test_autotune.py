"""Tuning table (cnn_tune_set / cnn_tune_get) and the tuner's class buckets;
no GPU needed: the table lives on the host."""

import ctypes
import os

import pytest

from commonnnclustering_amd import _lib, autotune


@pytest.fixture
def table():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libcnn_hip.so not built")
    L = _lib.lib()
    L.cnn_tune_clear()
    yield L
    L.cnn_tune_clear()
    autotune.load_cache()


def test_defaults_and_lookup(table):
    d = autotune.current(3, 10**6, "f32_band")
    assert d == {"mfma_col_chunk": 64, "wide_col_chunk": 32, "emit_waves": 8, "root_passes": 2,
                 "words_min_fill": 8}
    autotune.set_entry((2, 3), (1000, 2000), "f32_band", dict(d, root_passes=3))
    assert autotune.current(3, 1500, "f32_band")["root_passes"] == 3
    assert autotune.current(3, 2000, "f32_band")["root_passes"] == 2      # n_hi exclusive
    assert autotune.current(4, 1500, "f32_band")["root_passes"] == 2      # other d
    assert autotune.current(3, 1500, "f64")["root_passes"] == 2           # other precision
    autotune.set_entry((1, 8), (0, 1 << 20), "f32_band", dict(d, root_passes=1))
    assert autotune.current(3, 1500, "f32_band")["root_passes"] == 1      # later entries win


@pytest.mark.parametrize("bad", [dict(mfma_col_chunk=4), dict(emit_waves=6), dict(root_passes=9),
                                 dict(words_min_fill=-1), dict(wide_col_chunk=0)])
def test_invalid_entries_rejected(table, bad):
    p = dict(autotune.current(3, 100, "f64"), **bad)
    t = _lib.TuneParams(**p)
    assert table.cnn_tune_set(1, 3, 0, 100, 0, ctypes.byref(t)) == _lib.CNN_ERR_INVALID
    assert table.cnn_tune_set(3, 1, 0, 100, 0, ctypes.byref(_lib.TuneParams(**autotune.current(3, 100, "f64")))) \
        == _lib.CNN_ERR_INVALID


def test_buckets_cover_every_class():
    for d in (1, 2, 3, 5, 8, 9, 16, 40, 64, 100, 1000):
        for n in (1, 100, 20000, 200000, 10**6, 2**30):
            db, nb = autotune.bucket_of(d, n)
            assert db[0] <= d <= db[1] and nb[0] <= n < nb[1]


def test_cache_round_trip(table, tmp_path):
    p = tmp_path / "cache.json"
    import json

    json.dump({"entries": [{"d_lo": 16, "d_hi": 32, "n_lo": 0, "n_hi": 1 << 14, "precision": "f32_band",
                            "params": {"mfma_col_chunk": 128, "wide_col_chunk": 16, "emit_waves": 4,
                                       "root_passes": 1, "words_min_fill": 4}}]}, open(p, "w"))
    assert autotune.load_cache(str(p)) == 1
    assert autotune.current(20, 100, "f32_band")["mfma_col_chunk"] == 128
    assert autotune.current(20, 100, "f64")["mfma_col_chunk"] == 64
