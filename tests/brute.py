"""Brute-force restatement of one random-forest tree, for tiny inputs.

Independent of the oracle's code path: no ping-pong buffers, no running sums,
no mark table.  Every node re-sorts its own sample set per feature with
Python's ``sorted`` on (value, id), scores every eligible position by the
weighted Gini impurity computed from class fractions with exact
``fractions.Fraction`` arithmetic, and picks the minimum of
(impurity, feature id, position).  The feature visiting order pi is an input
(taken from the oracle's permutation function, which is pinned separately by
its own tests), like the bootstrap ids.

Reference behaviour restated (CudaTree):
  candidate positions   cudatree/cuda_kernels/common_func.cu:1, bfs_module.cu:173
  weighted Gini         cudatree/cuda_kernels/dfs_module.cu:404-407
  threshold             cudatree/cuda_kernels/bfs_module.cu:399-400
  leaf label            cudatree/random_tree.py:50-68
  leaf conditions       cudatree/random_tree.py:587-634
  traversal / vote      cudatree/cuda_kernels/predict.cu:50-67, random_forest.py:358
"""
from collections import Counter
from fractions import Fraction

import numpy as np


def _gini_weighted(labels_left, labels_right):
    n = len(labels_left) + len(labels_right)
    total = Fraction(0)
    for side in (labels_left, labels_right):
        ns = len(side)
        g = Fraction(1) - sum(Fraction(c, ns) ** 2 for c in Counter(side).values())
        total += Fraction(ns, n) * g
    return total


def best_split(X, y, samples, features):
    """Best (impurity, feature, position) over the given features for the
    sample set, or None.  Positions index the (value, id)-sorted order."""
    best = None
    for f in features:
        order = sorted(samples, key=lambda i: (float(X[i, f]), i))
        vals = [float(X[i, f]) for i in order]
        labs = [int(y[i]) for i in order]
        for p in range(len(order) - 1):
            if not (vals[p] + 1e-5 < vals[p + 1]):
                continue
            g = _gini_weighted(labs[: p + 1], labs[p + 1:])
            cand = (g, f, p)
            if best is None or cand < best:
                best = cand
    return best


def build_tree(X, y, K, samples, max_features, perm_fn, min_samples_split=1, max_depth=0):
    """Pre-order node arrays of one tree grown on the sample id list.
    perm_fn(depth, start) -> feature visiting order for the node whose
    segment starts at `start` (its offset in the sorted rows)."""
    X = np.asarray(X, dtype=np.float32)
    nodes = []

    def first_label(s):
        return int(y[min(s, key=lambda i: (float(X[i, 0]), i))])

    def grow(s, start, depth):
        nid = len(nodes)
        node = {"feature": 0, "threshold": np.float32(0), "left": 0, "right": 0, "value": 0,
                "n_samples": len(s), "depth": depth}
        nodes.append(node)
        labels = set(int(y[i]) for i in s)
        if len(labels) == 1:
            node["value"] = labels.pop()
            return
        if len(s) < min_samples_split:
            node["value"] = first_label(s)
            return
        if max_depth > 0 and depth >= max_depth:
            cnt = Counter(int(y[i]) for i in s)
            node["value"] = min(cnt, key=lambda k: (-cnt[k], k))
            return
        evaluated = []
        for f in perm_fn(depth, start):
            if len(evaluated) == max_features:
                break
            vals = sorted(float(X[i, f]) for i in s)
            if vals[0] == vals[-1]:
                continue
            evaluated.append(int(f))
        res = best_split(X, y, s, evaluated)
        if res is None:
            node["value"] = first_label(s)
            return
        _, f, p = res
        order = sorted(s, key=lambda i: (float(X[i, f]), i))
        a, b = X[order[p], f], X[order[p + 1], f]
        node["feature"] = f
        node["threshold"] = np.float32(np.float32(a + b) * np.float32(0.5))
        left = order[: p + 1]
        right = order[p + 1:]
        node["left"] = len(nodes)
        grow(left, start, depth + 1)
        node["right"] = len(nodes)
        grow(right, start + len(left), depth + 1)

    grow(list(samples), 0, 0)
    out = {}
    for key, dt in (("feature", np.int32), ("threshold", np.float32), ("left", np.uint32),
                    ("right", np.uint32), ("value", np.uint32), ("n_samples", np.uint32),
                    ("depth", np.uint32)):
        out[key] = np.array([n[key] for n in nodes], dtype=dt)
    return out


def predict(forest, X, K):
    """Traverse every tree for every row; majority vote, lowest label on ties."""
    X = np.asarray(X, dtype=np.float32)
    off = forest["tree_off"]
    out = np.zeros(len(X), np.uint32)
    for r, x in enumerate(X):
        votes = [0] * K
        for t in range(len(off) - 1):
            base = int(off[t])
            node = 0
            while forest["left"][base + node] != 0:
                f = forest["feature"][base + node]
                node = int(forest["left"][base + node] if x[f] < forest["threshold"][base + node]
                           else forest["right"][base + node])
            votes[int(forest["value"][base + node])] += 1
        out[r] = max(range(K), key=lambda k: (votes[k], -k))
    return out
