"""Seeded synthetic datasets for the fit/predict workloads, plus the
scikit-learn-bundled sets the reference's tests use.

Every generator is a ``numpy.random.Generator(PCG64(seed))`` stream, so a
(name, seed) pair always gives the same float32 row-major matrix and integer
labels.  The reference loads real data instead (cudatree/datasource.py:9-75);
covtype needs the network there, so the 581,012 x 54 tabular workload is a
synthetic stand-in with covtype's shape, class mix and one-hot layout.
"""
import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def digits_like(seed=1000, n=1797, n_features=64, n_classes=10):
    """1,797 x 64 integer-valued features in [0, 16], 10 balanced classes:
    per-class centroids uniform on {0..16}, samples rounded N(centroid, 3^2)."""
    rng = _rng(seed)
    centroid = rng.integers(0, 17, size=(n_classes, n_features))
    y = rng.permutation(np.arange(n) % n_classes)
    x = np.clip(np.rint(centroid[y] + rng.normal(0.0, 3.0, size=(n, n_features))), 0, 16)
    return x.astype(np.float32), y.astype(np.int64)


COVTYPE_PROPORTIONS = np.array([0.365, 0.488, 0.062, 0.005, 0.016, 0.030, 0.035])


def covtype_like(seed=1001, n_train=581012, n_test=100000):
    """581,012 x 54 float32, 7 skewed classes.  Features 0-9 are
    class-conditional two-component Gaussian mixtures over [0, 4000];
    10-13 a one-hot group of 4, 14-53 a one-hot group of 40, with
    class-dependent category probabilities.  Returns (X, y, X_test, y_test)."""
    rng = _rng(seed)
    n = n_train + n_test
    K = len(COVTYPE_PROPORTIONS)
    p = COVTYPE_PROPORTIONS / COVTYPE_PROPORTIONS.sum()
    y = rng.choice(K, size=n, p=p)
    X = np.zeros((n, 54), dtype=np.float32)
    means = rng.uniform(0.0, 4000.0, size=(K, 10, 2))
    sigma = rng.uniform(0.05, 0.15, size=(K, 10, 2)) * 4000.0
    comp = rng.integers(0, 2, size=(n, 10))
    mu = means[y[:, None], np.arange(10)[None, :], comp]
    sd = sigma[y[:, None], np.arange(10)[None, :], comp]
    X[:, :10] = (mu + sd * rng.standard_normal((n, 10))).astype(np.float32)
    for lo, width in ((10, 4), (14, 40)):
        probs = rng.dirichlet(np.ones(width), size=K)
        cum = np.cumsum(probs, axis=1)
        u = rng.random(n)
        cat = (u[:, None] > cum[y]).sum(axis=1)
        cat = np.minimum(cat, width - 1)
        X[np.arange(n), lo + cat] = 1.0
    return X[:n_train], y[:n_train], X[n_train:], y[n_train:]


def gaussian_clusters(seed=1002, n=200000, n_features=256, n_classes=20, clusters=3):
    """Gaussian mixture: `clusters` per class, centroids N(0, 2^2), sigma 1."""
    rng = _rng(seed)
    y = rng.integers(0, n_classes, size=n)
    cl = rng.integers(0, clusters, size=n)
    centroids = rng.normal(0.0, 2.0, size=(n_classes, clusters, n_features)).astype(np.float32)
    X = centroids[y, cl] + rng.standard_normal((n, n_features), dtype=np.float32)
    return X.astype(np.float32), y


def zipf_quantized(seed=1003, n=1000000, n_features=64, n_classes=100, s=1.1, levels=256):
    """Many classes with Zipf(s) frequencies and features quantised to
    `levels` uniform levels (many ties)."""
    rng = _rng(seed)
    w = 1.0 / np.arange(1, n_classes + 1) ** s
    y = rng.choice(n_classes, size=n, p=w / w.sum())
    centroids = rng.normal(0.0, 1.0, size=(n_classes, n_features)).astype(np.float32)
    X = centroids[y] + rng.standard_normal((n, n_features), dtype=np.float32)
    lo, hi = X.min(axis=0), X.max(axis=0)
    q = np.floor((X - lo) / (hi - lo) * (levels - 1) + 0.5)
    return q.astype(np.float32), y


def wide_sparse_signal(seed=1004, n=20000, n_features=4096, n_classes=5, informative=64):
    """Wide set: only the first `informative` features carry a 0.5 sigma
    class shift, the rest are N(0, 1) noise."""
    rng = _rng(seed)
    y = rng.integers(0, n_classes, size=n)
    X = rng.standard_normal((n, n_features), dtype=np.float32)
    shift = rng.choice([-0.5, 0.5], size=(n_classes, informative)).astype(np.float32)
    X[:, :informative] += shift[y]
    return X, y


def bootstrap_ids(n, n_trees, seed):
    """Per-tree in-bag sets: the sorted unique ids of n uniform draws from
    [0, n), without multiplicity, as the reference's
    np.unique(np.random.randint(0, n, n)) (cudatree/random_forest.py:123-126).
    Tree t draws from PCG64 seeded with [seed, t]."""
    out = []
    for t in range(n_trees):
        rng = np.random.Generator(np.random.PCG64([seed, t]))
        out.append(np.unique(rng.integers(0, n, size=n)).astype(np.uint32))
    return out


def load_sklearn(name):
    """The datasets of the reference's memorize tests (cudatree/datasource.py:12-23)."""
    from sklearn import datasets as skd

    if name == "digits":
        ds = skd.load_digits()
        return ds.data.astype(np.float32), ds.target
    if name == "iris":
        ds = skd.load_iris()
        return ds.data.astype(np.float32), ds.target
    if name == "diabetes":
        ds = skd.load_diabetes()
        return ds.data.astype(np.float32), (ds.target > 140).astype(np.int64)
    raise ValueError("unknown dataset %r" % name)


def tiny_case(seed):
    """A small random problem for exhaustive checks: ties, near-ties closer
    than 1e-5, constant features and uneven classes.  Returns a dict."""
    rng = _rng(seed)
    n = int(rng.integers(2, 48))
    F = int(rng.integers(1, 7))
    K = int(rng.integers(2, 5))
    kind = int(rng.integers(0, 4))
    if kind == 0:      # few distinct integer levels: many exact ties
        X = rng.integers(0, 4, size=(n, F)).astype(np.float32)
    elif kind == 1:    # continuous
        X = rng.normal(0, 1, size=(n, F)).astype(np.float32)
    elif kind == 2:    # near-ties below and above the 1e-5 candidate gap
        base = rng.integers(0, 3, size=(n, F)).astype(np.float64)
        jitter = rng.choice([0.0, 4e-6, 9e-6, 2e-5], size=(n, F))
        X = (base + jitter).astype(np.float32)
    else:              # one constant feature among the others
        X = rng.integers(0, 6, size=(n, F)).astype(np.float32)
        X[:, int(rng.integers(0, F))] = 3.0
    y = rng.integers(0, K, size=n)
    boot = bool(rng.integers(0, 2))
    return {
        "X": X, "y": y, "K": K,
        "m": int(rng.integers(1, F + 1)),
        "mss": int(rng.choice([1, 1, 2, 5])),
        "max_depth": int(rng.choice([0, 0, 1, 3])),
        "identity": bool(rng.integers(0, 4) == 0),
        "seed": int(rng.integers(0, 2**63)),
        "boot": boot,
        "n_trees": int(rng.integers(1, 4)),
    }
