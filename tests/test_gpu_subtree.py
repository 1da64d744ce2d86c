"""Small subtrees grown by single waves (needs an MI355X): for every
threshold -- none, every live child after the first level, and sizes in
between -- the forest equals the oracle's bit for bit, across the paths the
waves take (row-free splits, K > 64 with LDS class counters, 16- and 32-bit
entries, min_samples_split and max_depth leaves)."""
import numpy as np
import pytest

import helpers
import oracle
from cudatree_amd import _lib, datasets

pytestmark = pytest.mark.gpu

EVERY = 1 << 22  # every live child after the roots' level


@pytest.fixture(scope="module")
def ctx():
    return _lib.default_context(0)


@pytest.fixture(scope="module")
def orc():
    return oracle.get()


def fit(ctx, case, subtree_max, all_rows=0, hist_layout=0, idx_bits=0):
    ds = _lib.Dataset(ctx, case["X"], case["y"], case["K"], idx_bits=idx_bits)
    o = _lib.default_opts()
    o.seed = case["seed"]
    o.max_features = case["m"]
    o.min_samples_split = case["mss"]
    o.max_depth = case["max_depth"]
    o.identity_perm = int(case["identity"])
    o.all_rows = all_rows
    o.hist_layout = hist_layout
    o.subtree_max = subtree_max
    f = _lib.Forest.fit(ctx, ds, case["n_trees"], o, helpers.case_boot_ids(case))
    return f.export()


@pytest.mark.parametrize("sub", [0, 1, 8, 64, EVERY])
def test_tiny_cases(ctx, orc, sub):
    for s in range(60):
        case = datasets.tiny_case(s)
        print("subtree tiny sub=%d seed=%d" % (sub, s), flush=True)
        d = helpers.forest_diff(fit(ctx, case, sub), helpers.oracle_forest(case, orc))
        assert d is None, "seed %d: %s" % (s, d)


@pytest.mark.parametrize("sub,all_rows,idx_bits", [(512, 0, 0), (2048, 0, 32), (EVERY, 0, 0), (2048, 1, 0)])
def test_tabular(ctx, orc, sub, all_rows, idx_bits):
    X, y, _, _ = datasets.covtype_like(seed=3, n_train=20000, n_test=1)
    case = {"X": X, "y": y, "K": 7, "m": 5, "mss": 1, "max_depth": 0, "identity": False, "seed": 17,
            "boot": True, "n_trees": 4}
    print("subtree tabular sub=%d all_rows=%d idx_bits=%d" % (sub, all_rows, idx_bits), flush=True)
    d = helpers.forest_diff(fit(ctx, case, sub, all_rows=all_rows, idx_bits=idx_bits),
                            helpers.oracle_forest(case, orc))
    assert d is None, d


@pytest.mark.parametrize("mss,max_depth", [(5, 0), (1, 6), (20, 9)])
def test_leaf_rules(ctx, orc, mss, max_depth):
    X, y, _, _ = datasets.covtype_like(seed=8, n_train=8000, n_test=1)
    case = {"X": X, "y": y, "K": 7, "m": 7, "mss": mss, "max_depth": max_depth, "identity": False, "seed": 5,
            "boot": True, "n_trees": 5}
    print("subtree leaf rules mss=%d max_depth=%d" % (mss, max_depth), flush=True)
    d = helpers.forest_diff(fit(ctx, case, EVERY), helpers.oracle_forest(case, orc))
    assert d is None, d


@pytest.mark.parametrize("K,layout", [(300, 0), (20, 2)])
def test_many_classes(ctx, orc, K, layout):
    rng = np.random.default_rng(K)
    X = rng.normal(size=(3000, 12)).astype(np.float32)
    X[:, 2] = np.round(X[:, 2] * 2)
    y = rng.integers(0, K, size=3000).astype(np.uint32)
    y[:K] = np.arange(K)
    case = {"X": X, "y": y, "K": K, "m": 4, "mss": 1, "max_depth": 0, "identity": False, "seed": 9,
            "boot": True, "n_trees": 3}
    print("subtree K=%d layout=%d" % (K, layout), flush=True)
    d = helpers.forest_diff(fit(ctx, case, EVERY, hist_layout=layout), helpers.oracle_forest(case, orc))
    assert d is None, d


def test_zipf_32bit_entries(ctx, orc):
    X, y = datasets.zipf_quantized(n=80000)
    K = int(y.max()) + 1
    case = {"X": X, "y": y, "K": K, "m": 8, "mss": 1, "max_depth": 0, "identity": False, "seed": 1003,
            "boot": True, "n_trees": 2}
    print("subtree zipf K=%d" % K, flush=True)
    d = helpers.forest_diff(fit(ctx, case, 4096, idx_bits=32), helpers.oracle_forest(case, orc))
    assert d is None, d
