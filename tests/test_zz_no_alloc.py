"""The allocation-free cost model prediction (av1h_cm_predict_noalloc),
through the C ABI: its predictions equal those of av1h_cm_predict, the entry
point it is a variant of, bit for bit (the same kernel on the same inputs),
on two calls in a row, the second with fewer superblocks; a call over the
scratch array's capacity is refused."""
import ctypes

import numpy as np
import pytest

import av1ongpu_amd as pkg

pytestmark = pytest.mark.gpu

F = ctypes.POINTER(ctypes.c_float)


def _ptr(a):
    return a.ctypes.data_as(F)


def _noalloc(lib, model, X):
    X = np.ascontiguousarray(X, np.float32)
    out = np.full(X.shape[0], np.nan, np.float32)
    rc = lib.av1h_cm_predict_noalloc(model.handle, _ptr(X), X.shape[0], _ptr(out))
    return rc, out


def test_noalloc_prediction_matches_predict():
    lib = pkg.hip_lib()
    model = pkg.CostModel(0, seed=7)
    rng = np.random.default_rng(3)
    X = rng.gamma(2.0, 3.0, size=(3000, pkg.CM_FEATURES)).astype(np.float32)
    y = (X[:, :4].sum(axis=1) * 5.0 + rng.normal(0, 1, 3000)).astype(np.float32)
    model.train(X, y, epochs=2, lr=1e-3, seed=1)
    for n in (3000, 700):
        want = model.predict(X[:n])
        rc, got = _noalloc(lib, model, X[:n])
        assert rc == pkg.AV1H_OK
        assert np.array_equal(got, want), f"{n} superblocks: {np.abs(got - want).max()} apart"
    big = np.zeros((4097, pkg.CM_FEATURES), np.float32)
    rc, _ = _noalloc(lib, model, big)
    assert rc == -2  # AV1H_ERR_UNSUPPORTED: beyond the scratch array
    model.close()
