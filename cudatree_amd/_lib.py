"""ctypes binding of libcudatree_amd.so (the C-ABI of include/cudatree_amd.h).

The library is the only compute path: if it is missing or cannot reach a
GPU, every call raises.  There is no CPU fallback.
"""
import atexit
import ctypes
import itertools
import os
import threading
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcudatree_amd.so")

CTA_MODE_BATCHED = 0
CTA_MODE_HYBRID = 1
CTA_MODE_PER_NODE = 2
MODES = {"batched": CTA_MODE_BATCHED, "hybrid": CTA_MODE_HYBRID, "per_node": CTA_MODE_PER_NODE}
CTA_HIST_AUTO, CTA_HIST_REGS, CTA_HIST_LDS = 0, 1, 2

MAX_SAMPLES = 4194303
MAX_FEATURES = 65535
MAX_CLASSES = 1024

_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p
# void (*)(void *): the level hook of cta_ctx_set_level_hook
LEVEL_HOOK = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


class FitOpts(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("max_features", ctypes.c_int32),
        ("min_samples_split", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("identity_perm", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("bfs_threshold", ctypes.c_int64),
        ("scan_chunk", ctypes.c_int32),
        ("resh_chunk", ctypes.c_int32),
        ("trees_per_batch", ctypes.c_int32),
        ("resh_global", ctypes.c_int32),
        ("bootstrap", ctypes.c_int32),
        ("all_rows", ctypes.c_int32),
        ("lanes", ctypes.c_int32),
        ("fill_global", ctypes.c_int32),
        ("hist_layout", ctypes.c_int32),
        ("subtree_max", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 1),
    ]


class ServerOpts(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("coalesce", ctypes.c_int32),
        ("max_batch_rows", ctypes.c_int64),
        ("hbm_budget", ctypes.c_uint64),
        ("fit_workers", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("max_batch_work", ctypes.c_int64),
        ("fit_yield_us", ctypes.c_int32),
        ("split_rows", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 2),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("predict_requests", ctypes.c_uint64),
        ("predict_rows", ctypes.c_uint64),
        ("predict_launches", ctypes.c_uint64),
        ("fit_jobs", ctypes.c_uint64),
        ("fit_rejected", ctypes.c_uint64),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("mean_ms", ctypes.c_double),
        ("rows_per_s", ctypes.c_double),
        ("hbm_budget", ctypes.c_uint64),
        ("hbm_in_use", ctypes.c_uint64),
        ("hbm_peak", ctypes.c_uint64),
        ("resident_forests", ctypes.c_int32),
        ("queued_predict", ctypes.c_int32),
        ("queued_fit", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("fit_mem_measured_max", ctypes.c_uint64),
        ("fit_mem_ratio_max", ctypes.c_double),
    ]


class KernelStat(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 32),
        ("launches", ctypes.c_uint64),
        ("total_ms", ctypes.c_double),
        ("alg_bytes", ctypes.c_uint64),
    ]


# name -> (restype, argtypes); every entry is declared in include/cudatree_amd.h
SIGNATURES = {
    "cta_last_error": (ctypes.c_char_p, []),
    "cta_version": (ctypes.c_int, []),
    "cta_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "cta_ctx_destroy": (ctypes.c_int, [_vp]),
    "cta_ctx_synchronize": (ctypes.c_int, [_vp]),
    "cta_ctx_mem_info": (ctypes.c_int, [_vp, _u64p, _u64p]),
    "cta_dataset_create": (ctypes.c_int, [_vp, _f32p, _u32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.POINTER(_vp)]),
    "cta_dataset_create_device": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.c_int32, ctypes.POINTER(_vp)]),
    "cta_dataset_create_keys": (ctypes.c_int, [_vp, _f32p, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32, _i32p, _u64p, ctypes.POINTER(_vp)]),
    "cta_dataset_destroy": (ctypes.c_int, [_vp]),
    "cta_dataset_info": (ctypes.c_int, [_vp, _i32p, _i32p, _i32p, _i32p]),
    "cta_dataset_sorted": (ctypes.c_int, [_vp, _u32p]),
    "cta_fit_opts_default": (None, [ctypes.POINTER(FitOpts)]),
    "cta_fit": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, _u32p, _i64p, ctypes.POINTER(FitOpts),
                               ctypes.POINTER(_vp)]),
    "cta_forest_destroy": (ctypes.c_int, [_vp]),
    "cta_forest_info": (ctypes.c_int, [_vp, _i32p, _i64p, _i32p, _i32p]),
    "cta_forest_export": (ctypes.c_int, [_vp, _i64p, _i32p, _f32p, _u32p, _u32p, _u32p, _u32p, _u32p]),
    "cta_forest_import": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _i64p, _i32p, _f32p,
                                         _u32p, _u32p, _u32p, ctypes.POINTER(_vp)]),
    "cta_forest_save": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, _u64p]),
    "cta_forest_load": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "cta_predict": (ctypes.c_int, [_vp, _vp, _f32p, ctypes.c_int64, ctypes.c_int32, _u32p, _u32p]),
    "cta_predict_device": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int32, _vp, _vp]),
    "cta_predict_device_noalloc": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int32, _vp]),
    "cta_ctx_release_workspace": (ctypes.c_int, [_vp]),
    "cta_ctx_set_level_hook": (ctypes.c_int, [_vp, LEVEL_HOOK, _vp]),
    "cta_ctx_workspace_bytes": (ctypes.c_int, [_vp, _u64p]),
    "cta_ctx_set_workspace_limit": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "cta_fit_estimate": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.POINTER(FitOpts), _u64p, _u64p]),
    "cta_server_opts_default": (None, [ctypes.POINTER(ServerOpts)]),
    "cta_server_last_error": (ctypes.c_char_p, []),
    "cta_server_create": (ctypes.c_int, [ctypes.POINTER(ServerOpts), ctypes.POINTER(_vp)]),
    "cta_server_destroy": (ctypes.c_int, [_vp]),
    "cta_server_add_forest": (ctypes.c_int, [_vp, _vp, _i32p]),
    "cta_server_drop_forest": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "cta_server_get_forest": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(_vp)]),
    "cta_server_submit_predict": (ctypes.c_int, [_vp, ctypes.c_int32, _f32p, ctypes.c_int64, ctypes.c_int32, _u32p,
                                                 _u64p]),
    "cta_server_submit_predict_many": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_f32p),
                                                      _i64p, ctypes.c_int32, ctypes.POINTER(_u32p), _u64p]),
    "cta_server_submit_fit": (ctypes.c_int, [_vp, _f32p, _u32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.POINTER(FitOpts), _u64p]),
    "cta_server_wait": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_int64, _i32p, _i32p]),
    "cta_server_wait_many": (ctypes.c_int, [_vp, ctypes.c_int32, _u64p, ctypes.c_int64, _i32p, _i32p]),
    "cta_server_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ServerStats)]),
    "cta_server_reset_stats": (ctypes.c_int, [_vp]),
    "cta_ctx_set_profiling": (ctypes.c_int, [_vp, ctypes.c_int]),
    "cta_ctx_reset_stats": (ctypes.c_int, [_vp]),
    "cta_ctx_kernel_stats": (ctypes.c_int, [_vp, ctypes.POINTER(KernelStat), ctypes.c_int32, _i32p]),
    "cta_k_split_scan": (ctypes.c_int, [_vp, _vp, _u32p, ctypes.c_uint32, _i32p, ctypes.c_int32, ctypes.c_int32,
                                        _i32p, _u32p, _u64p, _u64p]),
    "cta_k_reshuffle": (ctypes.c_int, [_vp, _vp, _u32p, ctypes.c_uint32, _u8p, ctypes.c_int32, ctypes.c_int32,
                                       _u32p]),
}


class CtaError(RuntimeError):
    """A failed library call: `.code` is the CTA_E* status."""

    def __init__(self, code, msg):
        super().__init__("cudatree_amd error %d: %s" % (code, msg))
        self.code = code


_lib = None
_lock = threading.Lock()

# Open handles, closed at interpreter exit in dependency order (datasets,
# forests and servers before the contexts they use) while the interpreter
# and the HIP runtime are still whole.
_open = {"dataset": weakref.WeakSet(), "forest": weakref.WeakSet(), "server": weakref.WeakSet()}


@atexit.register
def _close_all():
    for kind in ("server", "forest", "dataset"):
        for obj in list(_open[kind]):
            try:
                obj.close()
            except Exception:
                pass
    for c in list(_ctx.values()):
        try:
            c.close()
        except Exception:
            pass


def load():
    """The loaded library; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError("libcudatree_amd.so not built at %s (run __graft_entry__.build() or "
                                  "make -C cudatree_amd/csrc)" % LIB_PATH)
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(rc):
    if rc != 0:
        msg = load().cta_last_error()
        raise CtaError(rc, msg.decode() if msg else "")


def ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Context:
    """One HIP device with its own stream (cta_ctx)."""

    def __init__(self, device=0):
        self.lib = load()
        h = _vp()
        check(self.lib.cta_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.lib.cta_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(self.lib.cta_ctx_synchronize(self.h))

    def mem_info(self):
        f, t = ctypes.c_uint64(), ctypes.c_uint64()
        check(self.lib.cta_ctx_mem_info(self.h, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def set_profiling(self, on):
        check(self.lib.cta_ctx_set_profiling(self.h, int(bool(on))))

    def reset_stats(self):
        check(self.lib.cta_ctx_reset_stats(self.h))

    def release_workspace(self):
        check(self.lib.cta_ctx_release_workspace(self.h))

    def set_level_hook(self, fn):
        """fn() runs on the fitting thread after each group of batched
        levels (None: no hook)."""
        self._hook = LEVEL_HOOK(lambda _arg: fn()) if fn is not None else None
        check(self.lib.cta_ctx_set_level_hook(self.h, self._hook if fn is not None else LEVEL_HOOK(), None))

    def kernel_stats(self):
        n = ctypes.c_int32()
        check(self.lib.cta_ctx_kernel_stats(self.h, None, 0, ctypes.byref(n)))
        arr = (KernelStat * max(n.value, 1))()
        check(self.lib.cta_ctx_kernel_stats(self.h, arr, n.value, ctypes.byref(n)))
        return {arr[i].name.decode(): {"launches": arr[i].launches, "ms": arr[i].total_ms,
                                       "alg_bytes": arr[i].alg_bytes} for i in range(n.value)}


_ctx = {}


def default_context(device=0):
    with _lock:
        c = _ctx.get(device)
    if c is None:
        c = Context(device)
        with _lock:
            _ctx.setdefault(device, c)
            c = _ctx[device]
    return c


class Dataset:
    """Training set resident on the device, presorted (cta_dataset)."""

    def __init__(self, ctx, X, y, K, idx_bits=0):
        self.ctx = ctx
        self.lib = ctx.lib
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.uint32)
        if X.ndim != 2 or y.ndim != 1 or len(y) != X.shape[0]:
            raise ValueError("X must be (N, F) and y (N,)")
        N, F = X.shape
        h = _vp()
        check(self.lib.cta_dataset_create(ctx.h, ptr(X, _f32p), ptr(y, _u32p), N, F, K, idx_bits, ctypes.byref(h)))
        self.h = h
        _open["dataset"].add(self)
        self.N, self.F, self.K = N, F, K
        b = ctypes.c_int32()
        check(self.lib.cta_dataset_info(h, None, None, None, ctypes.byref(b)))
        self.idx_bits = b.value

    @classmethod
    def from_labels(cls, ctx, X, target, idx_bits=0):
        """Training set with the class values of target compacted on the
        device (cta_dataset_create_keys): returns (dataset, classes), classes
        the distinct values of target in ascending order, a sample's label
        the index of its value.  None for a dtype class_keys cannot key."""
        ck = class_keys(target)
        if ck is None:
            return None
        keys, bits, decode = ck
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.ndim != 2 or keys.ndim != 1 or len(keys) != X.shape[0]:
            raise ValueError("X must be (N, F) and target (N,)")
        N, F = X.shape
        self = cls.__new__(cls)
        self.ctx, self.lib = ctx, ctx.lib
        h, K = _vp(), ctypes.c_int32()
        classes = np.zeros(MAX_CLASSES, np.uint64)
        check(self.lib.cta_dataset_create_keys(ctx.h, ptr(X, _f32p), keys.ctypes.data_as(_vp), bits, N, F, idx_bits,
                                               ctypes.byref(K), ptr(classes, _u64p), ctypes.byref(h)))
        self.h = h
        _open["dataset"].add(self)
        self.N, self.F, self.K = N, F, K.value
        b = ctypes.c_int32()
        check(self.lib.cta_dataset_info(h, None, None, None, ctypes.byref(b)))
        self.idx_bits = b.value
        return self, decode(classes[:K.value])

    def sorted_rows(self):
        out = np.zeros((self.F, self.N), np.uint32)
        check(self.lib.cta_dataset_sorted(self.h, ptr(out, _u32p)))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.cta_dataset_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def class_keys(target):
    """Order-preserving unsigned keys of numeric class values for
    cta_dataset_create_keys: (keys, key_bits, decode) with decode(keys)
    giving the class values back in target's dtype, or None for a dtype
    without such keys (strings, objects).  Equal values get equal keys
    (-0.0 and +0.0 included); NaN is refused, as np.unique would give every
    NaN a class of its own."""
    y = np.ascontiguousarray(target)
    dt = y.dtype
    if dt == np.bool_ or (dt.kind == "u" and dt.itemsize <= 4):
        return y.astype(np.uint32), 32, lambda k: k.astype(np.uint32).astype(dt)
    if dt.kind == "i" and dt.itemsize <= 4:
        return (y.astype(np.int32).view(np.uint32) ^ np.uint32(0x80000000)), 32, \
            lambda k: (k.astype(np.uint32) ^ np.uint32(0x80000000)).view(np.int32).astype(dt)
    if dt.kind == "u":
        return y.astype(np.uint64), 64, lambda k: k.astype(np.uint64).astype(dt)
    if dt.kind == "i":
        return (y.astype(np.int64).view(np.uint64) ^ np.uint64(1 << 63)), 64, \
            lambda k: (k.astype(np.uint64) ^ np.uint64(1 << 63)).view(np.int64).astype(dt)
    if dt.kind == "f" and dt.itemsize <= 8:
        v = y.astype(np.float64)
        if np.isnan(v).any():
            raise ValueError("class values contain NaN")
        u = (v + 0.0).view(np.uint64)  # -0.0 + 0.0 is +0.0
        sign = np.uint64(1 << 63)
        keys = np.where(u & sign, ~u, u | sign)

        def decode(k):
            k = k.astype(np.uint64)
            return np.where(k & sign, k ^ sign, ~k).view(np.float64).astype(dt)

        return keys, 64, decode
    return None


FOREST_FIELDS = ("feature", "threshold", "left", "right", "value", "n_samples", "depth")


class Forest:
    """A fitted or imported forest (cta_forest), resident for prediction."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.lib = ctx.lib
        self.h = h
        _open["forest"].add(self)
        T, n, F, K = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        check(self.lib.cta_forest_info(h, ctypes.byref(T), ctypes.byref(n), ctypes.byref(F), ctypes.byref(K)))
        self.n_trees, self.n_nodes, self.F, self.K = T.value, n.value, F.value, K.value

    @classmethod
    def fit(cls, ctx, ds, n_trees, opts, boot_ids=None):
        if boot_ids is not None:
            flat = np.ascontiguousarray(np.concatenate(boot_ids).astype(np.uint32))
            off = np.zeros(n_trees + 1, np.int64)
            off[1:] = np.cumsum([len(b) for b in boot_ids])
        else:
            flat = off = None
        h = _vp()
        check(ctx.lib.cta_fit(ctx.h, ds.h, n_trees, ptr(flat, _u32p), ptr(off, _i64p), ctypes.byref(opts),
                              ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_arrays(cls, ctx, arrays, F, K):
        off = np.ascontiguousarray(arrays["tree_off"], dtype=np.int64)
        T = len(off) - 1
        a = {k: np.ascontiguousarray(arrays[k], dtype=dt) for k, dt in
             (("feature", np.int32), ("threshold", np.float32), ("left", np.uint32), ("right", np.uint32),
              ("value", np.uint32))}
        h = _vp()
        check(ctx.lib.cta_forest_import(ctx.h, T, F, K, ptr(off, _i64p), ptr(a["feature"], _i32p),
                                        ptr(a["threshold"], _f32p), ptr(a["left"], _u32p), ptr(a["right"], _u32p),
                                        ptr(a["value"], _u32p), ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def load(cls, ctx, image):
        """Resident forest from a serialised image (cta_forest_load)."""
        buf = np.frombuffer(bytes(image), np.uint8)
        h = _vp()
        check(ctx.lib.cta_forest_load(ctx.h, buf.ctypes.data_as(_vp), buf.size, ctypes.byref(h)))
        return cls(ctx, h)

    def save(self):
        """Serialised image of the forest (cta_forest_save), as bytes."""
        size = ctypes.c_uint64()
        check(self.lib.cta_forest_save(self.h, None, 0, ctypes.byref(size)))
        buf = np.zeros(size.value, np.uint8)
        check(self.lib.cta_forest_save(self.h, buf.ctypes.data_as(_vp), size.value, ctypes.byref(size)))
        return buf.tobytes()

    def export(self):
        off = np.zeros(self.n_trees + 1, np.int64)
        n = self.n_nodes
        out = {"tree_off": off, "feature": np.zeros(n, np.int32), "threshold": np.zeros(n, np.float32)}
        for k in ("left", "right", "value", "n_samples", "depth"):
            out[k] = np.zeros(n, np.uint32)
        check(self.lib.cta_forest_export(self.h, ptr(off, _i64p), ptr(out["feature"], _i32p),
                                         ptr(out["threshold"], _f32p), ptr(out["left"], _u32p),
                                         ptr(out["right"], _u32p), ptr(out["value"], _u32p),
                                         ptr(out["n_samples"], _u32p), ptr(out["depth"], _u32p)))
        return out

    def predict(self, X, votes=False):
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.ndim != 2 or X.shape[1] != self.F:
            raise ValueError("X must be (R, %d)" % self.F)
        R = X.shape[0]
        labels = np.zeros(R, np.uint32)
        v = np.zeros((R, self.K), np.uint32) if votes else None
        check(self.lib.cta_predict(self.ctx.h, self.h, ptr(X, _f32p), R, self.F, ptr(labels, _u32p), ptr(v, _u32p)))
        return (labels, v) if votes else labels

    def close(self):
        if getattr(self, "h", None):
            self.lib.cta_forest_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fit_estimate(N, F, K, T, opts=None):
    """(dataset bytes, workspace bytes) the library estimates for a fit of T
    trees on an N x F set with K classes (host computation, no device)."""
    o = opts if opts is not None else default_opts()
    ds, ws = ctypes.c_uint64(), ctypes.c_uint64()
    check(load().cta_fit_estimate(N, F, K, T, ctypes.byref(o), ctypes.byref(ds), ctypes.byref(ws)))
    return ds.value, ws.value


def default_opts():
    o = FitOpts()
    load().cta_fit_opts_default(ctypes.byref(o))
    return o


def split_scan(ctx, ds, rows, features, chunk=0):
    """Device split scan of one node (test hook)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    feats = np.ascontiguousarray(features, dtype=np.int32)
    n = rows.shape[1]
    bf, bp, num, den = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64()
    check(ctx.lib.cta_k_split_scan(ctx.h, ds.h, ptr(rows, _u32p), n, ptr(feats, _i32p), len(feats), chunk,
                                   ctypes.byref(bf), ctypes.byref(bp), ctypes.byref(num), ctypes.byref(den)))
    if bf.value < 0:
        return None
    return bf.value, bp.value, num.value, den.value


def reshuffle(ctx, ds, rows, mark, chunk=0, global_bits=False):
    """Device stable partition of one node's rows (test hook)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    mark = np.ascontiguousarray(mark, dtype=np.uint8)
    out = np.zeros_like(rows)
    check(ctx.lib.cta_k_reshuffle(ctx.h, ds.h, ptr(rows, _u32p), rows.shape[1], ptr(mark, _u8p), chunk,
                                  int(bool(global_bits)), ptr(out, _u32p)))
    return out


def check_server(rc):
    if rc != 0:
        msg = load().cta_server_last_error()
        raise CtaError(rc, msg.decode() if msg else "")


class Server:
    """Serving loop (cta_server): resident forests, coalesced predict
    requests, fit jobs admitted by HBM accounting."""

    def __init__(self, device=0, coalesce=True, max_batch_rows=0, hbm_budget=0, fit_workers=1, max_batch_work=0,
                 fit_yield_us=0, split_rows=0):
        self.lib = load()
        o = ServerOpts()
        self.lib.cta_server_opts_default(ctypes.byref(o))
        o.device = device
        o.coalesce = int(bool(coalesce))
        if max_batch_rows:
            o.max_batch_rows = max_batch_rows
        if max_batch_work:
            o.max_batch_work = max_batch_work
        o.hbm_budget = hbm_budget
        o.fit_workers = fit_workers
        o.fit_yield_us = fit_yield_us
        o.split_rows = split_rows
        h = _vp()
        check_server(self.lib.cta_server_create(ctypes.byref(o), ctypes.byref(h)))
        self.h = h
        _open["server"].add(self)
        self._keep = {}
        self._lock = threading.Lock()

    def add_forest(self, forest):
        fid = ctypes.c_int32()
        check_server(self.lib.cta_server_add_forest(self.h, forest.h, ctypes.byref(fid)))
        return fid.value

    def drop_forest(self, fid):
        check_server(self.lib.cta_server_drop_forest(self.h, fid))

    def export_forest(self, fid):
        f = _vp()
        check_server(self.lib.cta_server_get_forest(self.h, fid, ctypes.byref(f)))
        borrowed = Forest.__new__(Forest)
        borrowed.lib = self.lib
        borrowed.h = f
        T, n, F, K = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        check(self.lib.cta_forest_info(f, ctypes.byref(T), ctypes.byref(n), ctypes.byref(F), ctypes.byref(K)))
        borrowed.n_trees, borrowed.n_nodes, borrowed.F, borrowed.K = T.value, n.value, F.value, K.value
        out = Forest.export(borrowed)
        borrowed.h = None  # borrowed: never destroyed from here
        return out

    def submit_predict(self, fid, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(X.shape[0], np.uint32)
        t = ctypes.c_uint64()
        check_server(self.lib.cta_server_submit_predict(self.h, fid, ptr(X, _f32p), X.shape[0], X.shape[1],
                                                        ptr(out, _u32p), ctypes.byref(t)))
        with self._lock:
            self._keep[t.value] = (X, out)
        return t.value

    def submit_many(self, fid, Xs):
        """Queues every array of Xs (rows, F) in one call; returns the tickets."""
        Xs = [np.ascontiguousarray(X, dtype=np.float32) for X in Xs]
        n = len(Xs)
        if n == 0:
            return []
        F = Xs[0].shape[1]
        outs = [np.zeros(X.shape[0], np.uint32) for X in Xs]
        xp = (_f32p * n)(*[ptr(X, _f32p) for X in Xs])
        lp = (_u32p * n)(*[ptr(o, _u32p) for o in outs])
        R = np.array([X.shape[0] for X in Xs], np.int64)
        t = np.zeros(n, np.uint64)
        check_server(self.lib.cta_server_submit_predict_many(self.h, fid, n, xp, ptr(R, _i64p), F, lp,
                                                             ptr(t, _u64p)))
        with self._lock:
            for k in range(n):
                self._keep[int(t[k])] = (Xs[k], outs[k])
        return [int(v) for v in t]

    def submit_rows(self, fid, X, R):
        """Queues len(R) requests whose rows lie back to back in X
        (sum(R) x F): request i is the R[i] rows after the first R[0] + ... +
        R[i-1].  Returns (tickets, labels): labels (sum(R) uint32) receives
        every request's labels at the same offsets.  Both arrays are held
        here until wait_many has reported every ticket of the call."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        R = np.ascontiguousarray(R, dtype=np.int64)
        n = len(R)
        if n == 0:
            return np.zeros(0, np.uint64), np.zeros(0, np.uint32)
        if R.min() < 1 or int(R.sum()) != X.shape[0]:
            raise ValueError("R must hold positive counts summing to X.shape[0]")
        F = X.shape[1]
        labels = np.zeros(X.shape[0], np.uint32)
        first = np.concatenate(([0], np.cumsum(R)[:-1])).astype(np.uint64)
        xp = np.uint64(X.ctypes.data) + first * np.uint64(F * 4)
        lp = np.uint64(labels.ctypes.data) + first * np.uint64(4)
        t = np.zeros(n, np.uint64)
        check_server(self.lib.cta_server_submit_predict_many(
            self.h, fid, n, xp.ctypes.data_as(ctypes.POINTER(_f32p)), ptr(R, _i64p), F,
            lp.ctypes.data_as(ctypes.POINTER(_u32p)), ptr(t, _u64p)))
        held = (X, labels)  # every ticket of the call holds the arrays
        with self._lock:
            self._keep.update(zip(t.tolist(), itertools.repeat(held)))
        return t, labels

    def wait_many(self, tickets, timeout_us=-1):
        """Statuses of tickets (uint64 array, in order) as wait reports them,
        with one deadline (-1: still pending)."""
        tickets = np.ascontiguousarray(tickets, dtype=np.uint64)
        st = np.zeros(len(tickets), np.int32)
        check_server(self.lib.cta_server_wait_many(self.h, len(tickets), ptr(tickets, _u64p), timeout_us,
                                                   ptr(st, _i32p), None))
        with self._lock:
            for k in tickets[st >= 0].tolist():
                self._keep.pop(k, None)
        return st

    def submit_fit(self, X, y, K, n_trees, opts):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.uint32)
        t = ctypes.c_uint64()
        check_server(self.lib.cta_server_submit_fit(self.h, ptr(X, _f32p), ptr(y, _u32p), X.shape[0], X.shape[1],
                                                    K, n_trees, ctypes.byref(opts), ctypes.byref(t)))
        return t.value

    def wait(self, ticket, timeout_us=-1):
        """(status, result): result is the labels of a predict ticket (for a
        submit_rows ticket, the call's whole label array) or the forest id of
        a fit ticket; status -1 means still pending."""
        st, fid = ctypes.c_int32(), ctypes.c_int32()
        check_server(self.lib.cta_server_wait(self.h, ticket, timeout_us, ctypes.byref(st), ctypes.byref(fid)))
        if st.value < 0:
            return -1, None
        with self._lock:
            kept = self._keep.pop(ticket, None)
        return st.value, (kept[1] if kept is not None else fid.value)

    def stats(self):
        s = ServerStats()
        check_server(self.lib.cta_server_get_stats(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_ if k != "reserved"}

    def reset_stats(self):
        check_server(self.lib.cta_server_reset_stats(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.cta_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- serialised forests ---------------------------------------------------------
# The byte image of cta_forest_save / cta_forest_load (include/cudatree_amd.h):
# a 48-byte header, tree offsets, then the seven node arrays.
FOREST_MAGIC = b"CTAFORST"
FOREST_FORMAT = 1
_HEADER = np.dtype([("magic", "S8"), ("version", "<u4"), ("header_bytes", "<u4"), ("T", "<i4"), ("F", "<i4"),
                    ("K", "<i4"), ("flags", "<u4"), ("nodes", "<i8"), ("reserved", "<u8")])
_IMAGE_ARRAYS = (("feature", "<i4"), ("threshold", "<f4"), ("left", "<u4"), ("right", "<u4"), ("value", "<u4"),
                 ("n_samples", "<u4"), ("depth", "<u4"))


def forest_to_bytes(arrays, F, K):
    """Byte image of exported forest arrays, as cta_forest_save writes it
    (host only; no device needed)."""
    off = np.ascontiguousarray(arrays["tree_off"], dtype="<i8")
    T, n = len(off) - 1, int(off[-1])
    h = np.zeros(1, _HEADER)
    h["magic"], h["version"], h["header_bytes"] = FOREST_MAGIC, FOREST_FORMAT, _HEADER.itemsize
    h["T"], h["F"], h["K"], h["nodes"] = T, F, K, n
    parts = [h.tobytes(), off.tobytes()]
    for k, dt in _IMAGE_ARRAYS:
        a = arrays.get(k)
        parts.append((np.zeros(n, dt) if a is None else np.ascontiguousarray(a, dtype=dt)).tobytes())
    return b"".join(parts)


def forest_from_bytes(image):
    """(arrays, F, K) of a byte image; ValueError for another format version
    or a malformed image (host only)."""
    image = bytes(image)
    if len(image) < _HEADER.itemsize:
        raise ValueError("forest image shorter than its header")
    h = np.frombuffer(image[:_HEADER.itemsize], _HEADER)[0]
    if h["magic"] != FOREST_MAGIC:
        raise ValueError("not a serialised forest (bad magic)")
    if int(h["version"]) != FOREST_FORMAT:
        raise ValueError("forest format version %d is not supported (this build reads version %d)"
                         % (int(h["version"]), FOREST_FORMAT))
    T, n = int(h["T"]), int(h["nodes"])
    if int(h["header_bytes"]) != _HEADER.itemsize or T <= 0 or n <= 0 or \
            len(image) != _HEADER.itemsize + 8 * (T + 1) + 28 * n:
        raise ValueError("corrupt forest image")
    p = _HEADER.itemsize
    out = {"tree_off": np.frombuffer(image, "<i8", T + 1, p).astype(np.int64)}
    p += 8 * (T + 1)
    for k, dt in _IMAGE_ARRAYS:
        out[k] = np.frombuffer(image, dt, n, p).copy()
        p += 4 * n
    if out["tree_off"][0] != 0 or out["tree_off"][-1] != n:
        raise ValueError("tree offsets do not match the node count")
    return out, int(h["F"]), int(h["K"])
