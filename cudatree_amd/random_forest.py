"""RandomForestClassifier with CudaTree's interface, on the MI355X library.

Mirrors cudatree/random_forest.py: the constructor arguments
(random_forest.py:45-51), ``fit(samples, target, bfs_threshold=None)``
(:292), ``predict`` (:339), ``predict_proba`` (:364), ``score`` (:379) and
pickling (:42-43).  Keyword-only extensions: ``max_depth``, ``random_state``,
``mode`` (batched / hybrid / per_node scheduling), ``device``, the scan /
reshuffle chunk sizes and ``autotune`` (time a probe of the problem to pick
the schedule).
"""
import math
import time

import numpy as np

from . import _lib


def _best_bfs_threshold(n_labels, n_samples, max_features):
    """The reference's default switch point, fitted on a GTX Titan
    (random_forest.py:278-289)."""
    t = int(4746 + 4 * n_labels + 0.0651 * n_samples - 75 * max_features)
    return max(min(t, 50000), 2000)


def _best_dtype(max_value):
    """Smallest unsigned dtype holding max_value (util.py:40-49)."""
    for dt in (np.uint8, np.uint16, np.uint32):
        if max_value <= np.iinfo(dt).max:
            return np.dtype(dt)
    return np.dtype(np.uint64)


class timer(object):
    """Prints the wall time of a block (util.py:51-60)."""

    def __init__(self, name):
        self.name = name

    def __enter__(self, *args):
        print("Running %s" % self.name)
        self.start_t = time.time()

    def __exit__(self, *args):
        print("Time for %s: %s" % (self.name, time.time() - self.start_t))


class RandomForestClassifier(object):
    """A random forest classifier grown on the GPU.

    Parameters follow cudatree.RandomForestClassifier:
      n_estimators (10), max_features (None -> ceil(sqrt(n_features))),
      min_samples_split (1), bootstrap (True), verbose, debug (features
      visited in id order at every node).
    """

    def __init__(self, n_estimators=10, max_features=None, min_samples_split=1, bootstrap=True, verbose=False,
                 debug=False, *, max_depth=None, random_state=None, mode="batched", device=0, scan_chunk=0,
                 resh_chunk=0, trees_per_batch=0, autotune=False):
        self.n_estimators = n_estimators
        self.max_features = max_features
        self.min_samples_split = min_samples_split
        self.bootstrap = bootstrap
        self.verbose = verbose
        self.debug = debug
        self.max_depth = max_depth
        self.random_state = random_state
        self.mode = mode
        self.device = device
        self.scan_chunk = scan_chunk
        self.resh_chunk = resh_chunk
        self.trees_per_batch = trees_per_batch
        self.autotune = autotune
        self.tuned_ = None
        self._forest = None
        self._arrays = None

    # -- fitting ------------------------------------------------------------
    def _seed(self):
        if self.random_state is None:
            return int(np.random.randint(0, 2**31 - 1)) * 2**32 + int(np.random.randint(0, 2**31 - 1))
        return int(self.random_state) & (2**64 - 1)

    def fit(self, samples, target, bfs_threshold=None):
        """Grow n_estimators trees on samples (n_samples, n_features) and
        target (n_samples,).  bfs_threshold: node size above which the
        'hybrid' mode grows nodes one at a time (default: the reference's
        regression, random_forest.py:278-289)."""
        samples = np.asarray(samples)
        target = np.asarray(target)
        if samples.ndim != 2 or target.ndim != 1 or samples.shape[0] != target.shape[0]:
            raise ValueError("samples must be (n_samples, n_features) and target (n_samples,)")
        N, F = samples.shape
        m = self.max_features
        if m is None:
            m = int(math.ceil(np.sqrt(F)))
        if not 1 <= m <= F:
            raise ValueError("max_features must be in [1, %d]" % F)
        ctx = _lib.default_context(self.device)
        # label compaction (random_forest.py:85-99): on the device for
        # numeric class values (the dataset is built in the same call), with
        # np.unique on the host for other dtypes and for the autotuner's probe
        ds = None
        if not self.autotune:
            made = _lib.Dataset.from_labels(ctx, samples, target)
            if made is not None:
                ds, self.compt_table = made
                y = None
        if ds is None:
            self.compt_table, y = np.unique(target, return_inverse=True)
        self.n_labels = len(self.compt_table)
        self.n_features = F
        self.dtype_labels = _best_dtype(self.n_labels)
        self.dtype_indices = max(_best_dtype(N), np.dtype(np.uint16))
        self.max_features_ = m
        self.bfs_threshold = _best_bfs_threshold(self.n_labels, N, m) if bfs_threshold is None else bfs_threshold
        seed = self._seed()
        self.seed_ = seed
        if ds is None:
            ds = _lib.Dataset(ctx, samples.astype(np.float32), y.astype(np.uint32), self.n_labels)
        opts = _lib.default_opts()
        # in-bag sets drawn on the device from (seed, tree): the sorted unique
        # ids of N uniform draws, as np.unique(randint(0, N, N)) in the
        # reference (random_forest.py:123-126)
        opts.bootstrap = int(bool(self.bootstrap))
        opts.seed = seed
        opts.max_features = m
        opts.min_samples_split = int(self.min_samples_split)
        opts.max_depth = int(self.max_depth or 0)
        opts.identity_perm = int(bool(self.debug))
        opts.mode = _lib.MODES[self.mode]
        opts.bfs_threshold = int(self.bfs_threshold)
        opts.scan_chunk = int(self.scan_chunk)
        opts.resh_chunk = int(self.resh_chunk)
        opts.trees_per_batch = int(self.trees_per_batch)
        if self.autotune:
            # schedule picked by timing a probe of this problem; every setting
            # grows the same trees
            from . import autotune as _autotune

            best, _ = _autotune.tune(samples, y.astype(np.uint32), self.n_labels, m, bootstrap=self.bootstrap,
                                     max_depth=self.max_depth, seed=seed, device=self.device)
            _autotune.apply(opts, best)
            self.tuned_ = best
        if self.verbose:
            print("bfs_threshold : %d; bootstrap : %r; min_samples_split : %d; mode : %s"
                  % (self.bfs_threshold, self.bootstrap, self.min_samples_split, self.mode))
            print("n_samples : %d; n_features : %d; n_labels : %d; max_features : %d"
                  % (N, F, self.n_labels, m))
        t0 = time.time()
        self._forest = _lib.Forest.fit(ctx, ds, self.n_estimators, opts, None)
        ds.close()
        self._arrays = None
        if self.verbose:
            print("fit %d trees in %.3f s" % (self.n_estimators, time.time() - t0))
        return self

    # -- trees ----------------------------------------------------------------
    def export_trees(self):
        """Canonical pre-order node arrays of all trees (dict of numpy arrays,
        tree t at [tree_off[t], tree_off[t+1]))."""
        if self._arrays is None:
            if self._forest is None:
                raise ValueError("forest is not fitted")
            self._arrays = self._forest.export()
        return self._arrays

    @property
    def trees_(self):
        a = self.export_trees()
        off = a["tree_off"]
        return [{k: a[k][off[t]:off[t + 1]] for k in _lib.FOREST_FIELDS} for t in range(len(off) - 1)]

    def _device_forest(self):
        if self._forest is None:
            if self._arrays is None:
                raise ValueError("forest is not fitted")
            ctx = _lib.default_context(self.device)
            self._forest = _lib.Forest.from_arrays(ctx, self._arrays, self.n_features, self.n_labels)
        return self._forest

    # -- prediction -------------------------------------------------------------
    def _votes(self, x):
        x = np.asarray(x, dtype=np.float32)
        if x.ndim != 2 or x.shape[1] != self.n_features:
            raise ValueError("x must be (n, %d)" % self.n_features)
        return self._device_forest().predict(x, votes=True)

    def predict(self, x):
        """Majority vote of the trees, mapped back to the original labels
        (random_forest.py:339-362; lowest label wins ties)."""
        labels, _ = self._votes(x)
        return self.compt_table[labels]

    def predict_proba(self, x):
        """Fraction of trees voting for each label (random_forest.py:364-377)."""
        _, votes = self._votes(x)
        return votes.astype(np.float64) / float(self.n_estimators)

    def score(self, X, Y):
        return np.mean(self.predict(X) == Y)

    # -- pickling -------------------------------------------------------------
    # Pickled state: the constructor and fit parameters, the label table
    # (compt_table, the original class values) and the trees as the versioned
    # byte image of cta_forest_save.  The reference pickled bare tree arrays
    # with no version and restored a forest that could not predict
    # (random_forest.py:17-23, 42-43).
    STATE_VERSION = 1

    def __getstate__(self):
        state = {k: v for k, v in self.__dict__.items() if k not in ("_forest", "_arrays")}
        state["state_version"] = self.STATE_VERSION
        if self._forest is not None or self._arrays is not None:
            state["forest_image"] = _lib.forest_to_bytes(self.export_trees(), self.n_features, self.n_labels)
        return state

    def __setstate__(self, state):
        version = state.get("state_version")
        if version != self.STATE_VERSION:
            raise ValueError("pickled RandomForestClassifier has state version %r; this build reads version %d"
                             % (version, self.STATE_VERSION))
        state = dict(state)
        image = state.pop("forest_image", None)
        state.pop("state_version")
        self.__dict__.update(state)
        self._forest = None
        self._arrays = None
        if image is not None:
            arrays, F, K = _lib.forest_from_bytes(image)
            if F != self.n_features or K != self.n_labels or len(arrays["tree_off"]) - 1 != self.n_estimators:
                raise ValueError("pickled forest image does not match the classifier's shape")
            self._arrays = arrays
