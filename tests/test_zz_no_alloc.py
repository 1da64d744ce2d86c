"""The allocation-free predict entry point (needs an MI355X).

cta_predict_device_noalloc keeps its vote matrix in a fixed per-GPU device
buffer; its labels must equal cta_predict_device's (and the oracle's) on two
calls in a row, the second with fewer rows, and sizes beyond the buffer are
refused before any device work.  Device buffers come from the library's own
HIP runtime (helpers.DeviceArray).
"""
import ctypes

import numpy as np
import pytest

import oracle
from cudatree_amd import _lib, datasets
from helpers import DeviceArray

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def _predict_device(lib, ctx, forest, Xd, R, F, out):
    _lib.check(lib.cta_predict_device(ctx.h, forest.h, Xd.ptr, R, F, out.ptr, None))


def _predict_noalloc(lib, ctx, forest, Xd, R, F, out):
    return lib.cta_predict_device_noalloc(ctx.h, forest.h, Xd.ptr, R, F, out.ptr)


def test_noalloc_predict_matches_predict_device():
    lib = _lib.load()
    ctx = _lib.default_context(0)
    X, y = datasets.digits_like(seed=1000, n=1797)
    K = int(y.max()) + 1
    ds = _lib.Dataset(ctx, X, y, K)
    o = _lib.default_opts()
    o.seed = 99
    o.max_features = 8
    o.bootstrap = 1
    forest = _lib.Forest.fit(ctx, ds, 20, o, None)
    want_oracle = oracle.get().predict(forest.export(), X, K)
    rng = np.random.default_rng(5)
    F = X.shape[1]
    for R in (1797, 333):  # the second call smaller: stale votes of the first must not leak
        rows = rng.choice(len(X), size=R, replace=False)
        Xd = DeviceArray.from_host(X[rows].astype(np.float32))
        ref = DeviceArray.from_host(np.zeros(R, np.uint32))
        got = DeviceArray.from_host(np.full(R, 0xFFFFFFFF, np.uint32))
        _predict_device(lib, ctx, forest, Xd, R, F, ref)
        rc = _predict_noalloc(lib, ctx, forest, Xd, R, F, got)
        assert rc == 0, lib.cta_last_error()
        g = got.host()
        assert np.array_equal(g, ref.host())
        assert np.array_equal(g, want_oracle[rows])
        for d in (Xd, ref, got):
            d.free()
    forest.close()
    ds.close()


def test_noalloc_predict_rejects_oversize():
    lib = _lib.load()
    ctx = _lib.default_context(0)
    X, y = datasets.digits_like(seed=1001, n=400)
    K = int(y.max()) + 1
    F = X.shape[1]
    ds = _lib.Dataset(ctx, X, y, K)
    o = _lib.default_opts()
    o.seed = 7
    o.max_features = 8
    forest = _lib.Forest.fit(ctx, ds, 2, o, None)
    limit = 1 << 20
    R = limit // K + 1  # R x K just above the buffer: refused on the host
    Xd = DeviceArray((R, F), np.float32)
    out = DeviceArray(R, np.uint32)
    rc = _predict_noalloc(lib, ctx, forest, Xd, R, F, out)
    assert rc == 4 and b"vote words" in lib.cta_last_error()
    Xd.free()
    out.free()
    # the largest allowed size still runs and agrees with cta_predict_device
    R = limit // K
    Xd = DeviceArray.from_host(np.resize(X, (R, F)).astype(np.float32))
    out = DeviceArray(R, np.uint32)
    ref = DeviceArray(R, np.uint32)
    assert _predict_noalloc(lib, ctx, forest, Xd, R, F, out) == 0, lib.cta_last_error()
    _predict_device(lib, ctx, forest, Xd, R, F, ref)
    assert np.array_equal(out.host(), ref.host())
    for d in (Xd, out, ref):
        d.free()
    forest.close()
    ds.close()
