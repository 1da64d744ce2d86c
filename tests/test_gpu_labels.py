"""Label compaction on the device (needs an MI355X): cta_dataset_create_keys
gives the classes and labels of np.unique(return_inverse=True)
(random_forest.py:85-99); a forest grown from it equals the one grown from
host-compacted labels, and the oracle's."""
import numpy as np
import pytest

import helpers
import oracle
from cudatree_amd import _lib, datasets

pytestmark = pytest.mark.gpu


def _fit(ctx, ds, T, seed):
    o = _lib.default_opts()
    o.seed = seed
    o.max_features = 6
    o.bootstrap = 1
    return _lib.Forest.fit(ctx, ds, T, o, None).export()


@pytest.mark.parametrize("case", ["int64_gaps", "float32", "k300", "tiny", "one_class", "uint8"])
def test_device_compaction_matches_np_unique(case):
    rng = np.random.default_rng(17)
    N = {"tiny": 900, "k300": 20000}.get(case, 70000)
    X = rng.normal(size=(N, 12)).astype(np.float32)
    X[:, 3] = np.round(X[:, 3])  # ties
    if case == "int64_gaps":
        vals = np.array([-(2**62), -7, 0, 5, 2**40, 2**62], np.int64)
        t = vals[rng.integers(0, len(vals), size=N)]
    elif case == "float32":
        vals = np.array([-2.5, -0.0, 0.0, 1e-3, 7.25, np.inf], np.float32)
        t = vals[rng.integers(0, len(vals), size=N)]
    elif case == "k300":
        t = (rng.integers(0, 300, size=N) * 7 - 1000).astype(np.int32)
    elif case == "tiny":
        t = rng.integers(0, 4, size=N).astype(np.int16)
    elif case == "one_class":
        t = np.full(N, 9, np.int64)
    else:
        t = rng.integers(0, 200, size=N).astype(np.uint8)
    want_c, want_y = np.unique(t, return_inverse=True)
    K = len(want_c)
    ctx = _lib.default_context(0)
    print("labels %s N=%d K=%d" % (case, N, K), flush=True)
    ds, classes = _lib.Dataset.from_labels(ctx, X, t)
    assert ds.K == K and classes.dtype == t.dtype and np.array_equal(classes, want_c)
    got = _fit(ctx, ds, 3, 41)
    ref_ds = _lib.Dataset(ctx, X, want_y.astype(np.uint32), K)
    want = _fit(ctx, ref_ds, 3, 41)
    assert helpers.forest_diff(got, want) is None
    if case in ("int64_gaps", "tiny"):
        orc = oracle.get()
        ids = orc.bootstrap_forest_ids(41, 3, N)
        assert helpers.forest_diff(got, orc.fit(X, want_y.astype(np.uint32), K, 3, 6, boot_ids=ids, seed=41)) is None
    ds.close()
    ref_ds.close()


def test_more_classes_than_the_limit_are_refused():
    rng = np.random.default_rng(2)
    X = rng.normal(size=(5000, 4)).astype(np.float32)
    t = rng.integers(0, 1500, size=5000).astype(np.int64)
    assert len(np.unique(t)) > 1024
    ctx = _lib.default_context(0)
    with pytest.raises(_lib.CtaError) as e:
        _lib.Dataset.from_labels(ctx, X, t)
    assert e.value.code == 4


def test_facade_labels_on_device_match_host_path():
    X, y = datasets.digits_like(seed=1003, n=1200)
    labels = (y.astype(np.int64) * 1000 - 4000)  # sparse, negative class values
    import cudatree_amd

    a = cudatree_amd.RandomForestClassifier(n_estimators=5, random_state=3).fit(X, labels)
    b_classes = np.unique(labels)
    assert np.array_equal(a.compt_table, b_classes)
    pred = a.predict(X)
    assert pred.dtype == labels.dtype and set(np.unique(pred)) <= set(b_classes.tolist())
    assert np.mean(pred == labels) > 0.95
