"""ctypes binding of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP path in
``cudatree_amd``; only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this module.  The product path never
does: ``cudatree_amd`` fails loudly when its HIP library is missing instead of
falling back here.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)


class _Params(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("max_features", ctypes.c_int32),
        ("min_samples_split", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("identity_perm", ctypes.c_int32),
    ]


class _Forest(ctypes.Structure):
    _fields_ = [
        ("n_trees", ctypes.c_int32),
        ("tree_off", _i64p),
        ("feature", _i32p),
        ("threshold", _f32p),
        ("left", _u32p),
        ("right", _u32p),
        ("value", _u32p),
        ("n_samples", _u32p),
        ("depth", _u32p),
    ]


FIELDS = ("feature", "threshold", "left", "right", "value", "n_samples", "depth")
MAX_SAMPLES = 4194303


def build(mutants=False):
    """Compile liboracle.so (and the planted-mistake variants) with make."""
    targets = ["all"] + (["mutants"] if mutants else [])
    subprocess.run(["make", "-s", "-C", _HERE] + targets, check=True)


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    """One loaded oracle library (the real one or a planted-mistake build)."""

    def __init__(self, path=_LIB_PATH):
        if not os.path.exists(path):
            build(mutants="mutants" in path)
        lib = ctypes.CDLL(path)
        lib.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        lib.orc_philox4x32_10.restype = None
        lib.orc_feature_perm.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32, _i32p]
        lib.orc_feature_perm.restype = None
        lib.orc_bootstrap_ids.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32, _u32p]
        lib.orc_bootstrap_ids.restype = ctypes.c_int64
        lib.orc_presort.argtypes = [_f32p, ctypes.c_int32, ctypes.c_int32, _u32p]
        lib.orc_presort.restype = ctypes.c_int
        lib.orc_scan_feature.argtypes = [_f32p, _u32p, ctypes.c_uint32, ctypes.c_uint32,
                                         _u32p, _u64p, _u64p]
        lib.orc_scan_feature.restype = ctypes.c_int
        lib.orc_fit.argtypes = [_f32p, _u32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _u32p,
                                ctypes.c_int32, _u32p, _i64p, ctypes.POINTER(_Params),
                                ctypes.POINTER(_Forest)]
        lib.orc_fit.restype = ctypes.c_int
        lib.orc_fit_threads.argtypes = [_f32p, _u32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _u32p,
                                        ctypes.c_int32, _u32p, _i64p, ctypes.POINTER(_Params), ctypes.c_int32,
                                        ctypes.POINTER(_Forest)]
        lib.orc_fit_threads.restype = ctypes.c_int
        lib.orc_forest_free.argtypes = [ctypes.POINTER(_Forest)]
        lib.orc_forest_free.restype = None
        lib.orc_predict.argtypes = [ctypes.POINTER(_Forest), _f32p, ctypes.c_int64, ctypes.c_int32,
                                    ctypes.c_int32, _u32p, _u32p]
        lib.orc_predict.restype = ctypes.c_int
        self.lib = lib
        self.path = path

    # -- primitives -------------------------------------------------------
    def philox(self, ctr, key):
        c = np.ascontiguousarray(ctr, dtype=np.uint32)
        k = np.ascontiguousarray(key, dtype=np.uint32)
        out = np.zeros(4, np.uint32)
        self.lib.orc_philox4x32_10(_ptr(c, _u32p), _ptr(k, _u32p), _ptr(out, _u32p))
        return out

    def feature_perm(self, seed, tree, depth, start, F, identity=False):
        out = np.zeros(F, np.int32)
        self.lib.orc_feature_perm(seed, tree, depth, start, F, int(identity), _ptr(out, _i32p))
        return out

    def bootstrap_ids(self, seed, tree, N):
        out = np.zeros(N, np.uint32)
        n = self.lib.orc_bootstrap_ids(seed, tree, N, _ptr(out, _u32p))
        if n < 0:
            raise MemoryError("orc_bootstrap_ids")
        return out[:n].copy()

    def bootstrap_forest_ids(self, seed, n_trees, N):
        """In-bag ids of trees 0..n_trees-1 under the canonical generator."""
        return [self.bootstrap_ids(seed, t, N) for t in range(n_trees)]

    def presort(self, X_T):
        X_T = np.ascontiguousarray(X_T, dtype=np.float32)
        F, N = X_T.shape
        out = np.zeros((F, N), np.uint32)
        rc = self.lib.orc_presort(_ptr(X_T, _f32p), N, F, _ptr(out, _u32p))
        if rc:
            raise ValueError("orc_presort failed with status %d" % rc)
        return out

    def scan_feature(self, v, lab, K):
        v = np.ascontiguousarray(v, dtype=np.float32)
        lab = np.ascontiguousarray(lab, dtype=np.uint32)
        p = ctypes.c_uint32()
        num = ctypes.c_uint64()
        den = ctypes.c_uint64()
        ok = self.lib.orc_scan_feature(_ptr(v, _f32p), _ptr(lab, _u32p), len(v), K,
                                       ctypes.byref(p), ctypes.byref(num), ctypes.byref(den))
        if not ok:
            return None
        return int(p.value), int(num.value), int(den.value)

    # -- forest -----------------------------------------------------------
    def fit(self, X, y, K, n_trees, max_features, boot_ids=None, seed=0, min_samples_split=1,
            max_depth=0, identity_perm=False, sorted_idx=None, threads=1):
        """Fit on row-major X (N x F) and compact labels y; returns a dict of
        flat node arrays plus 'tree_off'.  boot_ids: list of per-tree sorted
        unique id arrays, or None for bootstrap=False."""
        X = np.asarray(X, dtype=np.float32)
        N, F = X.shape
        X_T = np.ascontiguousarray(X.T)
        y = np.ascontiguousarray(y, dtype=np.uint32)
        prm = _Params(seed, max_features, min_samples_split, max_depth, int(identity_perm))
        if boot_ids is not None:
            flat = np.ascontiguousarray(np.concatenate(boot_ids).astype(np.uint32))
            off = np.zeros(n_trees + 1, np.int64)
            off[1:] = np.cumsum([len(b) for b in boot_ids])
            bptr, optr = _ptr(flat, _u32p), _ptr(off, _i64p)
        else:
            bptr, optr = None, None
        sp = None
        if sorted_idx is not None:
            sorted_idx = np.ascontiguousarray(sorted_idx, dtype=np.uint32)
            sp = _ptr(sorted_idx, _u32p)
        fo = _Forest()
        rc = self.lib.orc_fit_threads(_ptr(X_T, _f32p), _ptr(y, _u32p), N, F, K, sp, n_trees, bptr, optr,
                                      ctypes.byref(prm), int(threads), ctypes.byref(fo))
        if rc:
            raise ValueError("orc_fit failed with status %d" % rc)
        try:
            total = fo.tree_off[n_trees]
            res = {"tree_off": np.ctypeslib.as_array(fo.tree_off, (n_trees + 1,)).copy()}
            dt = {"feature": np.int32, "threshold": np.float32}
            for name in FIELDS:
                arr = np.ctypeslib.as_array(getattr(fo, name), (max(total, 1),))[:total]
                res[name] = arr.astype(dt.get(name, np.uint32)).copy()
        finally:
            self.lib.orc_forest_free(ctypes.byref(fo))
        return res

    def predict(self, forest, X, K, return_votes=False):
        X = np.ascontiguousarray(X, dtype=np.float32)
        R, F = X.shape
        T = len(forest["tree_off"]) - 1
        keep = {k: np.ascontiguousarray(forest[k]) for k in FIELDS}
        keep["tree_off"] = np.ascontiguousarray(forest["tree_off"], dtype=np.int64)
        fo = _Forest(T, _ptr(keep["tree_off"], _i64p), _ptr(keep["feature"].astype(np.int32), _i32p),
                     _ptr(keep["threshold"].astype(np.float32), _f32p),
                     _ptr(keep["left"].astype(np.uint32), _u32p), _ptr(keep["right"].astype(np.uint32), _u32p),
                     _ptr(keep["value"].astype(np.uint32), _u32p), None, None)
        # keep converted arrays alive for the call
        conv = [keep["feature"].astype(np.int32), keep["threshold"].astype(np.float32),
                keep["left"].astype(np.uint32), keep["right"].astype(np.uint32),
                keep["value"].astype(np.uint32)]
        fo.feature = _ptr(conv[0], _i32p)
        fo.threshold = _ptr(conv[1], _f32p)
        fo.left = _ptr(conv[2], _u32p)
        fo.right = _ptr(conv[3], _u32p)
        fo.value = _ptr(conv[4], _u32p)
        out = np.zeros(R, np.uint32)
        votes = np.zeros((R, K), np.uint32) if return_votes else None
        rc = self.lib.orc_predict(ctypes.byref(fo), _ptr(X, _f32p), R, F, K, _ptr(out, _u32p),
                                  _ptr(votes, _u32p) if votes is not None else None)
        if rc:
            raise ValueError("orc_predict failed with status %d" % rc)
        return (out, votes) if return_votes else out


_default = None


def get():
    """The oracle library, built on first use."""
    global _default
    if _default is None:
        _default = Oracle()
    return _default
