"""Autotuner for the fit schedule.

Successor of the reference's offline threshold regression
(estimate_threshold.py:57-96, random_forest.py:278-289), which fitted the
best ``bfs_threshold`` for a GTX Titan.  Here every setting grows the same
trees (the kernels are exact and the export is canonical), so a setting is
chosen on time alone: a grid of schedules is timed on a probe of the real
problem (a few trees on the full training set) and the fastest is cached per
problem shape.

Tuned settings:
  mode / bfs_threshold  batched levels over all trees, or the reference's
                        per-tree hybrid with nodes above bfs_threshold grown
                        one at a time (bfs_threshold in 256..65536)
  scan_chunk            samples per split-scan task (tile width of the scan)
  resh_chunk            samples per small-node reshuffle task
  resh_global           side bitmap staged in LDS or read from global memory
  hist_layout           the split scan's class counts: lane k holding class
                        k in a register (K <= 64) or per-wave LDS counters
"""
import itertools
import time

import numpy as np

from . import _lib

_cache = {}

DEFAULT_GRID = {
    "scan_chunk": [512, 1024, 2048, 8192],
    "resh_chunk": [2048, 4096, 8192],
    "resh_global": [0, 1],
    "hist_layout": [_lib.CTA_HIST_REGS, _lib.CTA_HIST_LDS],
}
HYBRID_THRESHOLDS = [256, 1024, 4096, 16384, 65536]


def _signature(N, F, K, m, bootstrap, max_depth):
    return (N, F, K, m, bool(bootstrap), int(max_depth or 0))


def candidate_settings(grid=None, hybrid=True, K=None):
    """Settings to try: the product of the grid in batched mode plus the
    hybrid schedule at each bfs_threshold (default chunks).  With K > 64
    classes the register layout does not exist and the layout knob drops."""
    grid = dict(grid or DEFAULT_GRID)
    if K is not None and K > 64 and "hist_layout" in grid:
        grid["hist_layout"] = [_lib.CTA_HIST_LDS]
    keys = sorted(grid)
    out = []
    for vals in itertools.product(*(grid[k] for k in keys)):
        s = dict(zip(keys, vals))
        s["mode"] = "batched"
        out.append(s)
    if hybrid:
        for t in HYBRID_THRESHOLDS:
            out.append({"mode": "hybrid", "bfs_threshold": t})
    return out


def apply(opts, setting):
    for k, v in setting.items():
        if k == "mode":
            opts.mode = _lib.MODES[v]
        else:
            setattr(opts, k, int(v))
    return opts


def tune(X, y, K, max_features, bootstrap=True, max_depth=0, seed=0, probe_trees=4, grid=None, hybrid=True,
         repeats=2, device=0, use_cache=True):
    """Times every candidate setting on `probe_trees` trees of the problem and
    returns (best_setting, table).  table rows: (setting, seconds)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    N, F = X.shape
    sig = _signature(N, F, K, max_features, bootstrap, max_depth)
    if use_cache and sig in _cache:
        return _cache[sig]
    ctx = _lib.default_context(device)
    ds = _lib.Dataset(ctx, X, y, K)
    table = []
    for setting in candidate_settings(grid, hybrid, K):
        best = None
        for _ in range(repeats):
            o = _lib.default_opts()
            o.seed = seed
            o.max_features = max_features
            o.max_depth = int(max_depth or 0)
            o.bootstrap = int(bool(bootstrap))
            apply(o, setting)
            t = time.perf_counter()
            f = _lib.Forest.fit(ctx, ds, probe_trees, o, None)
            dt = time.perf_counter() - t
            f.close()
            best = dt if best is None else min(best, dt)
        table.append((setting, best))
    ds.close()
    table.sort(key=lambda r: r[1])
    result = (table[0][0], table)
    _cache[sig] = result
    return result
