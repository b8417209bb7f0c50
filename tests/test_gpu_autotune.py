"""The autotuner on the device: every candidate shape gives the oracle's
labels (the tuner itself asserts that all candidates agree), and the winner
is stored for the class and used by later fits."""

import numpy as np
import pytest

from commonnnclustering_amd import _lib, autotune
from gpu_util import gpu_fit
from oracle import c_oracle as C

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("d,precision", [(16, "f32_band"), (20, "f64"), (3, "f32_band")])
def test_tune_class_and_fit(cuda, d, precision):
    rng = np.random.default_rng(d)
    n = 6000
    centres = rng.uniform(-4, 4, size=(5, d))
    X = (centres[rng.integers(0, 5, size=n)] + rng.normal(size=(n, d)) * 0.8).astype(np.float32)
    r = 2.2 if d > 3 else 0.5
    c, m = 8, 10
    L = _lib.lib()
    L.cnn_tune_clear()
    try:
        e = autotune.tune(X, r, c, precision, member_cutoff=m, steps=1)
        assert autotune.current(d, n, precision) == e["params"]
        assert e["ms_tuned"] <= e["ms_default"]
        assert {t["param"] for t in e["trials"]} <= set(autotune.relevant(d, precision))
        got, _, _ = gpu_fit(X, r, c, member_cutoff=m, precision=precision)
        assert (got == C.fit(X, r, c, member_cutoff=m, use_grid=d <= 3)).all()
    finally:
        L.cnn_tune_clear()
        autotune.load_cache()
