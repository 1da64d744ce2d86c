"""Parity at the shapes of the benchmark workloads (needs an MI355X).

Full-size fits run every tree on the device; the first trees of the forest
(each a function of the seed and its own tree index only) are grown by the
oracle from the same canonical in-bag sets and compared bit for bit:
feature, threshold bits, children, leaf labels, sample counts and depths.
Smaller shapes of the same generators compare every tree.  The shapes cover
32-bit row entries with labels in the entry (N > 65,535), K > 64 (class
counts in LDS), 16-bit ids at 4,096 features, hundreds of continuous
features without value codes, and 256-level quantised features (8-bit value
codes, many ties).
"""
import numpy as np
import pytest

import helpers
import oracle
from cudatree_amd import _lib, datasets

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _lib.default_context(0)
    yield c
    c.release_workspace()  # the full-size fits grow the context's workspaces to tens of GB


@pytest.fixture(scope="module")
def orc():
    return oracle.get()


def device_fit(ctx, X, y, K, T, m, seed, mode="batched", bfs_threshold=0, hist_layout=0, scan_chunk=0):
    """T trees with the in-bag sets drawn on the device from (seed, tree)."""
    ds = _lib.Dataset(ctx, X, y, K)
    o = _lib.default_opts()
    o.seed = seed
    o.max_features = m
    o.bootstrap = 1
    o.mode = _lib.MODES[mode]
    o.bfs_threshold = bfs_threshold
    o.hist_layout = hist_layout
    o.scan_chunk = scan_chunk
    f = _lib.Forest.fit(ctx, ds, T, o, None)
    return f, ds


def oracle_trees(orc, X, y, K, nt, m, seed):
    """The first nt trees of the same forest, grown by the oracle."""
    ids = orc.bootstrap_forest_ids(seed, nt, len(y))
    return orc.fit(X, y, K, nt, m, boot_ids=ids, seed=seed)


def check_first_trees(orc, got, X, y, K, nt, m, seed):
    want = oracle_trees(orc, X, y, K, nt, m, seed)
    d = helpers.forest_diff(helpers.forest_slice(got, 0, nt), want)
    assert d is None, d
    return want


def test_covtype_full_size(ctx, orc):
    """581,012 x 54, 7 skewed classes, 50 trees, m = 7 (the fit of the
    headline figure): trees 0-1 against the oracle; held-out predictions of
    those two trees against the oracle's."""
    X, y, Xt, _ = datasets.covtype_like()
    print("covtype 581012x54 50 trees", flush=True)
    f, ds = device_fit(ctx, X, y, 7, 50, 7, 20240601)
    assert ds.idx_bits == 32
    got = f.export()
    want = check_first_trees(orc, got, X, y, 7, 2, 7, 20240601)
    two = _lib.Forest.from_arrays(ctx, helpers.forest_slice(got, 0, 2), 54, 7)
    assert np.array_equal(two.predict(Xt[:20000]), orc.predict(want, Xt[:20000], 7))


@pytest.mark.parametrize("mode,layout", [("batched", 0), ("hybrid", 2)])
def test_gaussian_clusters_shape(ctx, orc, mode, layout):
    """50,000 x 256, 20 classes of 3 Gaussian clusters, m = 16: every
    feature continuous (no value codes), 32-bit entries; the hybrid schedule
    (nodes above 4,096 samples one at a time) with LDS class counters."""
    X, y = datasets.gaussian_clusters(n=50000)
    print("gaussian 50000x256 mode=%s layout=%d" % (mode, layout), flush=True)
    f, _ = device_fit(ctx, X, y, 20, 3, 16, 77, mode=mode, bfs_threshold=4096, hist_layout=layout)
    check_first_trees(orc, f.export(), X, y, 20, 3, 16, 77)


def test_gaussian_clusters_full_size(ctx, orc):
    """200,000 x 256, 20 classes, 200 trees, m = 16: tree 0 against the
    oracle."""
    X, y = datasets.gaussian_clusters()
    print("gaussian 200000x256 200 trees", flush=True)
    f, _ = device_fit(ctx, X, y, 20, 200, 16, 1002)
    got = f.export()
    assert f.n_trees == 200
    check_first_trees(orc, got, X, y, 20, 1, 16, 1002)


@pytest.mark.parametrize("chunk", [0, 512])
def test_zipf_many_classes_shape(ctx, orc, chunk):
    """150,000 x 64, K = 100 with Zipf(1.1) frequencies, 256-level features:
    labels in bits 22-31 of 32-bit entries and class counts in LDS (K > 64),
    8-bit value codes with many ties."""
    X, y = datasets.zipf_quantized(n=150000)
    K = int(y.max()) + 1
    assert K > 64
    print("zipf 150000x64 K=%d chunk=%d" % (K, chunk), flush=True)
    f, ds = device_fit(ctx, X, y, K, 2, 8, 1003, scan_chunk=chunk)
    assert ds.idx_bits == 32
    check_first_trees(orc, f.export(), X, y, K, 2, 8, 1003)


def test_zipf_many_classes_full_size(ctx, orc):
    """1,000,000 x 64, K = 100, 100 trees, m = 8 (a 125 KB side bitmap per
    tree, near the LDS limit): tree 0 against the oracle."""
    X, y = datasets.zipf_quantized()
    K = int(y.max()) + 1
    print("zipf 1000000x64 K=%d 100 trees" % K, flush=True)
    f, _ = device_fit(ctx, X, y, K, 100, 8, 1003)
    assert f.n_trees == 100
    check_first_trees(orc, f.export(), X, y, K, 1, 8, 1003)


def test_wide_sparse_full_size(ctx, orc):
    """20,000 x 4,096, 5 classes, 64 informative features, 100 trees, m = 64,
    16-bit sample ids: tree 0 against the oracle."""
    X, y = datasets.wide_sparse_signal()
    print("wide 20000x4096 100 trees", flush=True)
    f, ds = device_fit(ctx, X, y, 5, 100, 64, 1004)
    assert ds.idx_bits == 16 and f.n_trees == 100
    check_first_trees(orc, f.export(), X, y, 5, 1, 64, 1004)
