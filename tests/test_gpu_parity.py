"""Bit-exact parity of the HIP path with the CPU oracle (needs an MI355X).

Every comparison is exact: integer arrays equal, thresholds equal bit for
bit, predictions equal.  Small chunk sizes force the multi-chunk scan and
reshuffle paths on small inputs.
"""
import numpy as np
import pytest

import helpers
import oracle
from cudatree_amd import _lib, datasets

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return _lib.default_context(0)


@pytest.fixture(scope="module")
def orc():
    return oracle.get()


def gpu_forest(ctx, case, mode="batched", scan_chunk=0, resh_chunk=0, idx_bits=0, tpb=0, resh_global=0, all_rows=0,
               fill_global=0, hist_layout=0, subtree_max=0):
    X, y, K = case["X"], case["y"], case["K"]
    ds = _lib.Dataset(ctx, X, y, K, idx_bits=idx_bits)
    o = _lib.default_opts()
    o.seed = case["seed"]
    o.max_features = case["m"]
    o.min_samples_split = case["mss"]
    o.max_depth = case["max_depth"]
    o.identity_perm = int(case["identity"])
    o.mode = _lib.MODES[mode]
    o.bfs_threshold = case.get("bfs_threshold", 0)
    o.scan_chunk = scan_chunk
    o.resh_chunk = resh_chunk
    o.trees_per_batch = tpb
    o.resh_global = resh_global
    o.all_rows = all_rows
    o.fill_global = fill_global
    o.hist_layout = hist_layout
    o.subtree_max = subtree_max
    f = _lib.Forest.fit(ctx, ds, case["n_trees"], o, helpers.case_boot_ids(case))
    return f, ds


# ---- presort -------------------------------------------------------------
@pytest.mark.parametrize("N,idx_bits", [(1, 0), (37, 0), (5000, 0), (65535, 0), (70001, 0), (3000, 32)])
def test_presort_stable(ctx, N, idx_bits):
    print("presort N=%d idx_bits=%d" % (N, idx_bits), flush=True)
    rng = np.random.default_rng(N)
    F = 9
    X = np.empty((N, F), np.float32)
    X[:, 0] = rng.integers(0, 7, N)
    X[:, 1] = rng.normal(size=N)
    X[:, 2] = rng.choice([-0.0, 0.0, 1.0, -1.0, np.inf, -np.inf, 1e-38, -3e38], size=N)
    X[:, 3] = 5.0
    X[:, 4] = -rng.exponential(size=N)
    # two distinct keys (one partition pass): bits differing high, across the
    # sign, only in the lowest bit, and with signed zeros tied
    X[:, 5] = rng.integers(0, 2, N)
    X[:, 6] = rng.choice([-1.5, 2.25], size=N)
    X[:, 7] = rng.choice(np.array([1.0, np.nextafter(np.float32(1.0), np.float32(2.0))], np.float32), size=N)
    X[:, 8] = rng.choice([-0.0, 0.0, 3.0], size=N)
    y = rng.integers(0, 3, N)
    ds = _lib.Dataset(ctx, X, y, 3, idx_bits=idx_bits)
    assert ds.idx_bits == (idx_bits or (16 if N <= 65535 else 32))
    got = ds.sorted_rows()
    for f in range(F):
        assert np.array_equal(got[f], np.argsort(X[:, f], kind="stable")), f


def test_dataset_rejects_nan_and_bad_labels(ctx):
    X = np.zeros((10, 2), np.float32)
    X[3, 1] = np.nan
    with pytest.raises(_lib.CtaError):
        _lib.Dataset(ctx, X, np.zeros(10), 2)
    with pytest.raises(_lib.CtaError):
        _lib.Dataset(ctx, np.zeros((10, 2), np.float32), np.full(10, 5), 2)


# ---- kernel hooks ----------------------------------------------------------
def _node_rows(X, samples):
    """Sorted rows of a node: each feature's samples ordered by (value, id)."""
    F = X.shape[1]
    return np.stack([np.array(sorted(samples, key=lambda i: (float(X[i, f]), i)), np.uint32) for f in range(F)])


def _oracle_node_best(orc, X, y, K, rows, feats):
    best = None
    for f in feats:
        r = orc.scan_feature(X[rows[f], f], y[rows[f]], K)
        if r is None:
            continue
        p, num, den = r
        key = (num, den, f, p)
        if best is None:
            best = key
        else:
            a = num * best[1]
            b = best[0] * den
            if a > b or (a == b and f < best[2]):
                best = key
    return None if best is None else (best[2], best[3], best[0], best[1])


@pytest.mark.parametrize("chunk", [0, 64, 128, 200])
def test_split_scan_hook(ctx, orc, chunk):
    rng = np.random.default_rng(chunk + 1)
    for trial in range(12):
        N = int(rng.integers(50, 3000))
        F = int(rng.integers(1, 9))
        K = int(rng.choice([2, 3, 7, 70, 300]))
        X = rng.integers(0, int(rng.choice([3, 50, 1000])), size=(N, F)).astype(np.float32)
        y = rng.integers(0, K, N)
        ds = _lib.Dataset(ctx, X, y, K)
        samples = sorted(rng.choice(N, size=int(rng.integers(2, N + 1)), replace=False).tolist())
        rows = _node_rows(X, samples)
        feats = rng.permutation(F)[: int(rng.integers(1, F + 1))]
        print("split_scan chunk=%d trial=%d N=%d F=%d K=%d n=%d" % (chunk, trial, N, F, K, len(samples)),
              flush=True)
        got = _lib.split_scan(ctx, ds, rows, feats, chunk=chunk)
        want = _oracle_node_best(orc, X, y, K, rows, feats)
        assert got == want, (trial, got, want)


@pytest.mark.parametrize("chunk,global_bits", [(0, False), (64, False), (100, False), (0, True)])
def test_reshuffle_hook(ctx, chunk, global_bits):
    """Small nodes (one wave per row group) and nodes of >= 4,096 samples
    (whole workgroups per row, several 4,096-element tiles)."""
    rng = np.random.default_rng(chunk + 11 + int(global_bits))
    for trial in range(10):
        N = int(rng.integers(10, 30000))
        F = int(rng.integers(1, 6))
        X = rng.normal(size=(N, F)).astype(np.float32)
        ds = _lib.Dataset(ctx, X, np.zeros(N, np.int64), 1)
        samples = np.sort(rng.choice(N, size=int(rng.integers(1, N + 1)), replace=False))
        rows = np.stack([samples[np.argsort(X[samples, f], kind="stable")] for f in range(F)]).astype(np.uint32)
        mark = (rng.random(N) < rng.random()).astype(np.uint8)
        print("reshuffle chunk=%d global=%d trial=%d N=%d n=%d" % (chunk, global_bits, trial, N, len(samples)),
              flush=True)
        got = _lib.reshuffle(ctx, ds, rows, mark, chunk=chunk, global_bits=global_bits)
        for f in range(F):
            r = rows[f]
            want = np.concatenate([r[mark[r] == 1], r[mark[r] == 0]])
            assert np.array_equal(got[f], want), (trial, f)


# ---- whole forests -----------------------------------------------------------
TINY = list(range(60))


@pytest.mark.parametrize("mode,chunk,all_rows", [("batched", 0, 0), ("batched", 64, 0), ("per_node", 64, 0),
                                                ("hybrid", 64, 0), ("batched", 0, 1), ("hybrid", 64, 1)])
def test_forest_tiny_cases(ctx, orc, mode, chunk, all_rows):
    """Sixty small problems; two-valued features without rows (all_rows=0)
    and with (all_rows=1) build the same trees."""
    for s in TINY:
        case = datasets.tiny_case(s)
        case["bfs_threshold"] = 8
        print("tiny forest mode=%s chunk=%d all_rows=%d seed=%d" % (mode, chunk, all_rows, s), flush=True)
        f, _ = gpu_forest(ctx, case, mode=mode, scan_chunk=chunk, resh_chunk=chunk, all_rows=all_rows)
        d = helpers.forest_diff(f.export(), helpers.oracle_forest(case, orc))
        assert d is None, "seed %d: %s" % (s, d)


@pytest.mark.parametrize("idx_bits", [0, 32])
def test_forest_many_classes_u16_labels(ctx, orc, idx_bits):
    """K = 300: 16-bit labels, with 16-bit sample ids (labels read from y)
    and with 32-bit row entries (labels in bits 22-31 of each entry, class
    counts in LDS)."""
    rng = np.random.default_rng(9)
    N, F, K = 3000, 6, 300
    X = rng.integers(0, 40, size=(N, F)).astype(np.float32)
    y = rng.integers(0, K, N)
    case = {"X": X, "y": y, "K": K, "m": 3, "mss": 1, "max_depth": 0, "identity": False, "seed": 5,
            "boot": True, "n_trees": 3}
    print("forest K=300 idx_bits=%d" % idx_bits, flush=True)
    f, ds = gpu_forest(ctx, case, scan_chunk=256, resh_chunk=256, idx_bits=idx_bits)
    assert ds.idx_bits == (idx_bits or 16)
    assert helpers.forest_diff(f.export(), helpers.oracle_forest(case, orc)) is None


@pytest.mark.parametrize("mode", ["batched", "per_node"])
def test_fill_global_bitmaps(ctx, orc, mode):
    """Side bitmaps built with global atomics (fill_global=1: the path of
    N / 8 bytes beyond LDS), with row-free two-valued features counted by
    the fill tasks, bootstrap, several chunks per node."""
    X, y, _, _ = datasets.covtype_like(seed=29, n_train=20000, n_test=1)
    case = {"X": X, "y": y, "K": 7, "m": 7, "mss": 1, "max_depth": 0, "identity": False, "seed": 31,
            "boot": True, "n_trees": 3 if mode == "batched" else 1}
    want = helpers.oracle_forest(case, orc)
    for chunk in (0, 512):
        print("fill_global mode=%s chunk=%d" % (mode, chunk), flush=True)
        f, _ = gpu_forest(ctx, case, mode=mode, scan_chunk=chunk, fill_global=1)
        assert helpers.forest_diff(f.export(), want) is None, chunk
    for s in TINY[:20]:
        tc = datasets.tiny_case(s)
        print("fill_global tiny seed=%d" % s, flush=True)
        f, _ = gpu_forest(ctx, tc, mode=mode, fill_global=1, scan_chunk=64, resh_chunk=64)
        assert helpers.forest_diff(f.export(), helpers.oracle_forest(tc, orc)) is None, s


def test_hist_layout_lds(ctx, orc):
    """Split-scan class counts in per-wave LDS counters for K <= 64 (the
    layout K > 64 always takes) build the same trees; CTA_HIST_REGS is
    refused for K > 64."""
    X, y, _, _ = datasets.covtype_like(seed=37, n_train=15000, n_test=1)
    case = {"X": X, "y": y, "K": 7, "m": 7, "mss": 1, "max_depth": 0, "identity": False, "seed": 3,
            "boot": True, "n_trees": 3}
    want = helpers.oracle_forest(case, orc)
    for layout, chunk in ((1, 0), (2, 0), (2, 256)):
        print("hist_layout=%d chunk=%d" % (layout, chunk), flush=True)
        f, _ = gpu_forest(ctx, case, hist_layout=layout, scan_chunk=chunk)
        assert helpers.forest_diff(f.export(), want) is None, (layout, chunk)
    big = {"X": X[:500], "y": np.arange(500) % 70, "K": 70, "m": 7, "mss": 1, "max_depth": 0, "identity": False,
           "seed": 3, "boot": False, "n_trees": 1}
    with pytest.raises(_lib.CtaError):
        gpu_forest(ctx, big, hist_layout=1)


# (1.0, 1.0000100135803223) is a candidate pair in double precision but not in
# float (1 + 1e-5f rounds up to the second value); (1.0, 1.000009) is no
# candidate in either.
GAP_A = np.float32(1.0)
GAP_B = np.float32(1.0000100135803223)
GAP_C = np.float32(1.000009)


@pytest.mark.parametrize("all_rows", [0, 1])
def test_candidate_gap_double_on_device(ctx, orc, all_rows):
    """The candidate gap test is done in double on the device as in the
    oracle: a two-valued feature {GAP_A, GAP_B} (bit code column; row-free
    unless all_rows), a feature whose levels include a pair closer than the
    gap (no code column: the float gap test over X_T) and a feature {GAP_A,
    GAP_C} that is never a candidate.  The labels follow GAP_B, so the trees
    split on the pair."""
    rng = np.random.default_rng(43)
    N = 5000
    X = np.empty((N, 4), np.float32)
    X[:, 0] = rng.normal(size=N)
    hi1 = rng.random(N) < 0.4
    X[:, 1] = np.where(hi1, GAP_B, GAP_A)
    lv = rng.integers(0, 4, N)
    X[:, 2] = np.array([0.5, 0.500004, GAP_A, GAP_B], np.float32)[lv]
    X[:, 3] = np.where(rng.random(N) < 0.5, GAP_A, GAP_C)
    y = (hi1.astype(int) + 2 * (lv == 3) + (rng.random(N) < 0.05)) % 4
    case = {"X": X, "y": y, "K": 4, "m": 2, "mss": 1, "max_depth": 0, "identity": False, "seed": 47,
            "boot": True, "n_trees": 4}
    print("gap pair all_rows=%d" % all_rows, flush=True)
    f, _ = gpu_forest(ctx, case, all_rows=all_rows)
    got = f.export()
    assert helpers.forest_diff(got, helpers.oracle_forest(case, orc)) is None
    split = got["left"] != 0
    mid = np.float32((GAP_A + GAP_B) * np.float32(0.5))
    assert np.any(split & (got["feature"] == 1)) and np.all(got["threshold"][split & (got["feature"] == 1)] == mid)
    assert not np.any(split & (got["feature"] == 3))


def test_forest_digits_like_config(ctx, orc):
    """1,797 x 64, 10 classes, bootstrap=False, m=8, 50 trees."""
    X, y = datasets.digits_like()
    case = {"X": X, "y": y, "K": 10, "m": 8, "mss": 1, "max_depth": 0, "identity": False, "seed": 2024,
            "boot": False, "n_trees": 50}
    print("digits-like 50 trees", flush=True)
    f, _ = gpu_forest(ctx, case)
    got = f.export()
    want = helpers.oracle_forest(case, orc)
    assert helpers.forest_diff(got, want) is None
    assert np.array_equal(f.predict(X), orc.predict(want, X, 10))
    assert np.array_equal(f.predict(X), y)   # memorised (bootstrap=False)


@pytest.mark.parametrize("idx_bits,all_rows", [(0, 0), (32, 0), (0, 1)])
def test_forest_tabular_u32_indices(ctx, orc, idx_bits, all_rows):
    """covtype-shaped data above 65,535 rows (32-bit indices), bootstrap; the
    44 one-hot features without rows and with."""
    X, y, _, _ = datasets.covtype_like(seed=7, n_train=70000, n_test=1)
    case = {"X": X, "y": y, "K": 7, "m": 7, "mss": 1, "max_depth": 0, "identity": False, "seed": 99,
            "boot": True, "n_trees": 3}
    print("tabular 70000x54 idx_bits=%d all_rows=%d" % (idx_bits, all_rows), flush=True)
    f, ds = gpu_forest(ctx, case, idx_bits=idx_bits, all_rows=all_rows)
    assert ds.idx_bits == 32
    got = f.export()
    assert helpers.forest_diff(got, helpers.oracle_forest(case, orc)) is None


@pytest.mark.parametrize("rowfree_first", [False, True])
def test_score_bound_keeps_ties(ctx, orc, rowfree_first):
    """Chunks of a large node are ruled out by a score bound against the
    candidates the node already has (its best row-free split, the feature's
    chunk boundaries).  A continuous feature and a two-valued one give the
    same perfect split of 65,536 samples, tied in score: the lower feature
    index wins, so the chunk holding the continuous feature's tied position
    is scanned, while its chunks far from the cut are ruled out."""
    n, cut = 65536, 20000
    rng = np.random.default_rng(41)
    v = rng.permutation(n).astype(np.float32)  # distinct values
    y = (v >= cut).astype(np.int64)
    two = (v >= cut).astype(np.float32)
    noise = rng.normal(size=n).astype(np.float32)
    X = np.stack([noise, two, v] if rowfree_first else [noise, v, two], axis=1)
    case = {"X": X, "y": y, "K": 2, "m": 3, "mss": 2, "max_depth": 0, "identity": True, "seed": 5, "boot": False,
            "n_trees": 2}
    print("score-bound tie rowfree_first=%d" % rowfree_first, flush=True)
    f, _ = gpu_forest(ctx, case)
    got = f.export()
    assert helpers.forest_diff(got, helpers.oracle_forest(case, orc)) is None
    assert got["feature"][0] == 1
    assert got["threshold"][0] == (np.float32(0.5) if rowfree_first else np.float32(19999.5))


@pytest.mark.parametrize("idx_bits,all_rows", [(0, 0), (32, 0), (0, 1)])
def test_value_code_columns(ctx, orc, idx_bits, all_rows):
    """Features of every value-code kind: binary (bit column), < 256 integer
    levels (8-bit), thousands of integer levels (16-bit), levels 4e-6 apart
    (closer than the 1e-5 candidate gap: no column, float gap test),
    continuous floats, and a feature holding -0.0, +0.0 and 1.0.  Multi-chunk
    scans read the codes past each chunk and batch boundary."""
    rng = np.random.default_rng(31)
    N = 6000
    signed_zero = np.zeros(N)
    signed_zero[rng.random(N) < 0.4] = -0.0
    signed_zero[rng.random(N) < 0.3] = 1.0
    cols = [rng.integers(0, 2, N).astype(np.float64),
            rng.integers(0, 200, N).astype(np.float64),
            rng.integers(0, 60000, N).astype(np.float64),
            rng.integers(0, 3, N) + np.where(rng.random(N) < 0.5, 0.0, 4e-6),
            rng.standard_normal(N),
            signed_zero]
    X = np.stack(cols, 1).astype(np.float32)
    assert np.signbit(X[:, 5]).any() and (X[:, 5] == 0).sum() > np.signbit(X[:, 5]).sum()
    y = ((X[:, 0] > 0).astype(int) + (X[:, 1] > 100) + (X[:, 3] > 1) + (X[:, 5] > 0.5)
         + (rng.random(N) < 0.1)) % 4
    case = {"X": X, "y": y, "K": 4, "m": 3, "mss": 1, "max_depth": 0, "identity": False, "seed": 41,
            "boot": True, "n_trees": 4}
    print("value codes idx_bits=%d all_rows=%d" % (idx_bits, all_rows), flush=True)
    f, _ = gpu_forest(ctx, case, idx_bits=idx_bits, scan_chunk=512, all_rows=all_rows)
    assert helpers.forest_diff(f.export(), helpers.oracle_forest(case, orc)) is None


def test_row_free_budget_split(ctx, orc):
    """Forty two-valued features at K = 16: the count kernels hold 32 of them
    without rows, the rest keep rows; feature 0 is two-valued and keeps its
    row.  Trees equal the oracle's."""
    rng = np.random.default_rng(77)
    N, K = 3000, 16
    X = np.concatenate([rng.integers(0, 2, size=(N, 40)).astype(np.float32) * rng.choice([1.0, 3.5], size=40),
                        rng.normal(size=(N, 2)).astype(np.float32)], axis=1).astype(np.float32)
    y = (X[:, :6].sum(1).astype(int) + (X[:, 40] > 0) * 3 + rng.integers(0, 3, N)) % K
    case = {"X": X, "y": y, "K": K, "m": 9, "mss": 1, "max_depth": 0, "identity": False, "seed": 23,
            "boot": True, "n_trees": 4}
    for all_rows in (0, 1):
        print("row-free budget all_rows=%d" % all_rows, flush=True)
        f, _ = gpu_forest(ctx, case, all_rows=all_rows)
        assert helpers.forest_diff(f.export(), helpers.oracle_forest(case, orc)) is None, all_rows


def test_modes_and_batching_identical(ctx, orc):
    X, y, _, _ = datasets.covtype_like(seed=3, n_train=20000, n_test=1)
    case = {"X": X, "y": y, "K": 7, "m": 5, "mss": 2, "max_depth": 0, "identity": False, "seed": 17,
            "boot": True, "n_trees": 4, "bfs_threshold": 3000}
    want = helpers.oracle_forest(case, orc)
    for mode, tpb, chunk, glob in [("batched", 0, 0, 0), ("batched", 1, 512, 0), ("batched", 3, 0, 1),
                                   ("hybrid", 0, 1024, 0), ("per_node", 0, 0, 0)]:
        print("modes: %s tpb=%d chunk=%d global=%d" % (mode, tpb, chunk, glob), flush=True)
        f, _ = gpu_forest(ctx, case, mode=mode, tpb=tpb, scan_chunk=chunk, resh_chunk=chunk, resh_global=glob)
        assert helpers.forest_diff(f.export(), want) is None, (mode, tpb, chunk, glob)


@pytest.mark.parametrize("tpb", [0, 1, 2])
def test_device_bootstrap_matches_oracle_generator(ctx, orc, tpb):
    """In-bag sets drawn on the device equal the oracle's canonical generator
    (keyed by the global tree index, whatever the batch split)."""
    X, y, _, _ = datasets.covtype_like(seed=13, n_train=12000, n_test=1)
    case = {"X": X, "y": y, "K": 7, "m": 7, "mss": 1, "max_depth": 0, "identity": False,
            "seed": 0x1234567890ABCDEF, "boot": False, "n_trees": 5}
    ds = _lib.Dataset(ctx, X, y, 7)
    o = _lib.default_opts()
    o.seed = case["seed"]
    o.max_features = 7
    o.trees_per_batch = tpb
    o.bootstrap = 1
    print("device bootstrap tpb=%d" % tpb, flush=True)
    f = _lib.Forest.fit(ctx, ds, 5, o, None)
    ids = orc.bootstrap_forest_ids(case["seed"], 5, len(y))
    want = orc.fit(X, y, 7, 5, 7, boot_ids=ids, seed=case["seed"])
    got = f.export()
    assert helpers.forest_diff(got, want) is None
    assert [int(got["n_samples"][got["tree_off"][t]]) for t in range(5)] == [len(i) for i in ids]


@pytest.mark.parametrize("T,tpb", [(20, 0), (40, 8)])
def test_two_lanes_match_one(ctx, orc, T, tpb):
    """Batches run on two streams at once (lanes=2, alternate batches) give
    the trees and the predictions of one stream: one batch per lane (predict
    table joined on the device) and five batches (joined on the host)."""
    X, y, Xt, _ = datasets.covtype_like(seed=23, n_train=20000, n_test=3000)
    ds = _lib.Dataset(ctx, X, y, 7)
    out = {}
    for lanes in (1, 2):
        o = _lib.default_opts()
        o.seed = 77
        o.max_features = 7
        o.bootstrap = 1
        o.trees_per_batch = tpb
        o.lanes = lanes
        print("lanes=%d T=%d tpb=%d" % (lanes, T, tpb), flush=True)
        f = _lib.Forest.fit(ctx, ds, T, o, None)
        out[lanes] = (f.export(), f.predict(Xt))
    assert helpers.forest_diff(out[2][0], out[1][0]) is None
    assert np.array_equal(out[2][1], out[1][1])
    if tpb == 0:
        ids = orc.bootstrap_forest_ids(77, T, len(y))
        want = orc.fit(X, y, 7, T, 7, boot_ids=ids, seed=77)
        assert helpers.forest_diff(out[2][0], want) is None


def test_level_hook_runs_between_level_groups(ctx, orc):
    """The context's level hook runs on the fitting thread between groups of
    levels; the trees do not change.  Every node stays in the levels here
    (subtree_max < 0), so there is a group per four depths."""
    X, y, _, _ = datasets.covtype_like(seed=21, n_train=6000, n_test=1)
    case = {"X": X, "y": y, "K": 7, "m": 7, "mss": 1, "max_depth": 0, "identity": False, "seed": 5,
            "boot": True, "n_trees": 3}
    calls = []
    ctx.set_level_hook(lambda: calls.append(1))
    try:
        print("level hook", flush=True)
        f, _ = gpu_forest(ctx, case, subtree_max=-1)
    finally:
        ctx.set_level_hook(None)
    got = f.export()
    assert helpers.forest_diff(got, helpers.oracle_forest(case, orc)) is None
    depth = int(got["depth"].max())
    assert len(calls) >= (depth + 3) // 4  # four levels per group


def test_max_depth_and_min_samples_split(ctx, orc):
    X, y, _, _ = datasets.covtype_like(seed=5, n_train=8000, n_test=1)
    for md, mss in [(1, 1), (4, 1), (6, 20), (0, 50)]:
        case = {"X": X, "y": y, "K": 7, "m": 7, "mss": mss, "max_depth": md, "identity": False, "seed": 1,
                "boot": True, "n_trees": 2}
        print("max_depth=%d mss=%d" % (md, mss), flush=True)
        f, _ = gpu_forest(ctx, case)
        assert helpers.forest_diff(f.export(), helpers.oracle_forest(case, orc)) is None


# ---- prediction ----------------------------------------------------------------
@pytest.mark.parametrize("n_trees", [7, 13])
def test_predict_matches_oracle_and_import(ctx, orc, n_trees):
    """Votes and labels equal the oracle's: tree by tree (7 trees), and with
    the trees split over the XCDs (13 trees, uneven classes)."""
    X, y, Xt, yt = datasets.covtype_like(seed=11, n_train=10000, n_test=5000)
    case = {"X": X, "y": y, "K": 7, "m": 7, "mss": 1, "max_depth": 0, "identity": False, "seed": 8,
            "boot": True, "n_trees": n_trees}
    print("predict n_trees=%d" % n_trees, flush=True)
    f, _ = gpu_forest(ctx, case)
    arrays = f.export()
    want = orc.predict(arrays, Xt, 7, return_votes=True)
    got = f.predict(Xt, votes=True)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    g = _lib.Forest.from_arrays(ctx, arrays, 54, 7)
    got2 = g.predict(Xt, votes=True)
    assert np.array_equal(got2[0], want[0]) and np.array_equal(got2[1], want[1])
    # a forest of oracle-built trees predicts identically through the device
    o_arr = helpers.oracle_forest(case, orc)
    h = _lib.Forest.from_arrays(ctx, o_arr, 54, 7)
    assert np.array_equal(h.predict(Xt), orc.predict(o_arr, Xt, 7))


def test_import_rejects_malformed(ctx):
    bad = {"tree_off": np.array([0, 3]), "feature": np.array([0, 0, 0]), "threshold": np.zeros(3, np.float32),
           "left": np.array([2, 0, 0]), "right": np.array([1, 0, 0]), "value": np.array([0, 1, 0])}
    with pytest.raises(_lib.CtaError):
        _lib.Forest.from_arrays(ctx, bad, 1, 2)


# ---- facade ----------------------------------------------------------------------
@pytest.mark.parametrize("name", ["digits", "iris", "diabetes"])
def test_facade_memorize(name):
    """The reference's memorize tests through the CudaTree-style API."""
    import cudatree_amd

    x, y = cudatree_amd.load_data(name)
    forest = cudatree_amd.RandomForestClassifier(n_estimators=6, bootstrap=False, random_state=3)
    forest.fit(x, y)
    assert np.array_equal(forest.predict(x), y)


def test_facade_pickle_and_proba():
    import pickle

    import cudatree_amd

    x, y = cudatree_amd.load_data("digits")
    labels = np.array(["c%d" % v for v in y])
    f = cudatree_amd.RandomForestClassifier(n_estimators=9, random_state=4).fit(x, labels)
    p = f.predict(x)
    assert p.dtype.kind == "U" and np.mean(p == labels) > 0.97
    g = pickle.loads(pickle.dumps(f))
    assert np.array_equal(g.predict(x), p)
    proba = g.predict_proba(x[:100])
    assert proba.shape == (100, 10) and np.allclose(proba.sum(axis=1), 1.0)
    assert np.array_equal(f.compt_table[np.argmax(proba, axis=1)], p[:100])
