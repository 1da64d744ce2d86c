"""Shared test helpers: brute-force forests and forest comparison."""
import numpy as np

import brute
from cudatree_amd import datasets

FIELDS = ("feature", "threshold", "left", "right", "value", "n_samples", "depth")


def case_boot_ids(case):
    if not case["boot"]:
        return None
    return datasets.bootstrap_ids(len(case["y"]), case["n_trees"], case["seed"] & 0xFFFF)


def brute_forest(case, orc):
    """Forest grown by the brute-force builder with the oracle's feature
    permutation as input."""
    X, y, K = case["X"], case["y"], case["K"]
    N, F = X.shape
    ids = case_boot_ids(case)
    trees = []
    for t in range(case["n_trees"]):
        samples = list(range(N)) if ids is None else [int(i) for i in ids[t]]

        def perm(depth, start, t=t):
            return orc.feature_perm(case["seed"], t, depth, start, F, case["identity"])

        trees.append(brute.build_tree(X, y, K, samples, case["m"], perm, case["mss"], case["max_depth"]))
    return concat_trees(trees)


def concat_trees(trees):
    out = {"tree_off": np.zeros(len(trees) + 1, np.int64)}
    out["tree_off"][1:] = np.cumsum([len(t["feature"]) for t in trees])
    for k in FIELDS:
        out[k] = np.concatenate([t[k] for t in trees])
    return out


def oracle_forest(case, orc):
    return orc.fit(case["X"], case["y"], case["K"], case["n_trees"], case["m"],
                   boot_ids=case_boot_ids(case), seed=case["seed"],
                   min_samples_split=case["mss"], max_depth=case["max_depth"],
                   identity_perm=case["identity"])


def forest_diff(a, b, fields=FIELDS):
    """First differing field as a readable string, or None if identical
    (thresholds compared bit for bit)."""
    if not np.array_equal(np.asarray(a["tree_off"]), np.asarray(b["tree_off"])):
        return "tree_off %s vs %s" % (a["tree_off"], b["tree_off"])
    for k in fields:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        if k == "threshold":
            x, y = x.astype(np.float32).view(np.uint32), y.astype(np.float32).view(np.uint32)
        if x.shape != y.shape or not np.array_equal(x, y):
            idx = np.nonzero(x != y)[0] if x.shape == y.shape else []
            return "%s differs (first at %s)" % (k, idx[:5])
    return None


def forest_slice(a, t0, t1):
    """Trees [t0, t1) of an exported forest (node arrays and tree offsets
    rebased to 0)."""
    off = np.asarray(a["tree_off"])
    b, e = int(off[t0]), int(off[t1])
    out = {"tree_off": off[t0:t1 + 1] - b}
    for k in FIELDS:
        if k in a:
            out[k] = np.asarray(a[k])[b:e]
    return out


class DeviceArray:
    """A device buffer from the HIP runtime the library itself links (no
    torch, whose bundled runtime is a second copy in the process), with
    host copies in and out.  For tests of the *_device entry points."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes

            from cudatree_amd import _lib

            _lib.load()  # the library pulls in its HIP runtime
            cls._hip = ctypes.CDLL("libamdhip64.so.7")
        return cls._hip

    def __init__(self, shape, dtype):
        import ctypes

        self.shape, self.dtype = tuple(np.atleast_1d(shape)), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        rc = self.hip().hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.nbytes, 1)))
        assert rc == 0, "hipMalloc failed (%d)" % rc

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        import ctypes

        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        rc = self.hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.nbytes), 1)
        assert rc == 0, "hipMemcpy to the device failed (%d)" % rc

    def host(self):
        import ctypes

        out = np.empty(self.shape, self.dtype)
        rc = self.hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes), 2)
        assert rc == 0, "hipMemcpy to the host failed (%d)" % rc
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None
