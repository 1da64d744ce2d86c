"""Pins the CPU oracle: known-answer vectors, closed forms, brute force on
tiny inputs, the reference's own memorize/accuracy properties, and planted
mistakes that those checks must catch."""
import os
import subprocess

import numpy as np
import pytest

import brute
import helpers
import oracle
from cudatree_amd import datasets


@pytest.fixture(scope="module")
def orc():
    return oracle.get()


# Random123 known-answer vectors for philox4x32-10 (kat_vectors file of the
# Random123 distribution, D. E. Shaw Research).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(orc, ctr, key, expect):
    assert tuple(int(v) for v in orc.philox(ctr, key)) == expect


@pytest.mark.parametrize("F", [1, 2, 3, 4, 5, 7, 16, 17, 54, 64, 65, 256, 4096])
def test_feature_perm_is_permutation(orc, F):
    for key in [(0, 0, 0, 0), (7, 3, 2, 100), (2**63 + 5, 199, 40, 123456)]:
        p = orc.feature_perm(*key, F)
        assert sorted(p.tolist()) == list(range(F))
        assert np.array_equal(p, orc.feature_perm(*key, F))      # pure function of the key
    assert np.array_equal(orc.feature_perm(5, 1, 2, 3, F, identity=True), np.arange(F))


def test_feature_perm_varies_with_every_key_part(orc):
    base = orc.feature_perm(11, 2, 3, 4, 54)
    for alt in [(12, 2, 3, 4), (11, 3, 3, 4), (11, 2, 4, 4), (11, 2, 3, 5)]:
        assert not np.array_equal(base, orc.feature_perm(*alt, 54))


def test_feature_perm_first_slot_roughly_uniform(orc):
    F = 54
    firsts = np.array([orc.feature_perm(1, 0, d, s, F)[0] for d in range(20) for s in range(200)])
    counts = np.bincount(firsts, minlength=F)
    assert counts.min() > 0.4 * len(firsts) / F and counts.max() < 1.8 * len(firsts) / F


@pytest.mark.parametrize("N", [1, 2, 7, 100, 5000, 70001])
def test_bootstrap_ids_invariants(orc, N):
    ids = orc.bootstrap_ids(123, 4, N)
    assert ids.dtype == np.uint32 and len(ids) >= 1
    assert np.all(np.diff(ids.astype(np.int64)) > 0) and ids[-1] < N
    assert np.array_equal(ids, orc.bootstrap_ids(123, 4, N))
    if N >= 100:
        # the in-bag fraction of N draws with replacement: 1 - (1 - 1/N)^N
        expect = 1 - (1 - 1.0 / N) ** N
        sd = np.sqrt(N * expect * (1 - expect))
        assert abs(len(ids) - N * expect) < 6 * sd + 2
        assert not np.array_equal(ids, orc.bootstrap_ids(123, 5, N)[: len(ids)])
        assert not np.array_equal(ids, orc.bootstrap_ids(124, 4, N)[: len(ids)])


def test_bootstrap_ids_uniform_coverage(orc):
    N, T = 1000, 400
    hits = np.zeros(N)
    for t in range(T):
        hits[orc.bootstrap_ids(9, t, N)] += 1
    p = 1 - (1 - 1.0 / N) ** N
    # every sample is in-bag in about p * T trees (binomial, 6 sigma)
    assert np.all(np.abs(hits - p * T) < 6 * np.sqrt(T * p * (1 - p)))
    assert abs(hits[: N // 2].mean() - hits[N // 2:].mean()) < 3.0


def test_presort_matches_stable_argsort(orc):
    rng = np.random.default_rng(3)
    X_T = rng.integers(0, 5, size=(6, 500)).astype(np.float32)
    X_T[2] = rng.normal(size=500)
    X_T[3, :10] = -0.0
    X_T[3, 10:20] = 0.0
    out = orc.presort(X_T)
    for f in range(6):
        assert np.array_equal(out[f], np.argsort(X_T[f], kind="stable"))


def test_presort_rejects_nan(orc):
    X_T = np.zeros((1, 4), np.float32)
    X_T[0, 2] = np.nan
    with pytest.raises(ValueError):
        orc.presort(X_T)


def test_scan_feature_matches_bruteforce(orc):
    rng = np.random.default_rng(5)
    for trial in range(400):
        n = int(rng.integers(1, 40))
        K = int(rng.integers(1, 5))
        v = np.sort(rng.choice([0.0, 1.0, 1.000004, 1.00002, 2.0, 3.5], size=n)).astype(np.float32)
        lab = rng.integers(0, K, size=n)
        got = orc.scan_feature(v, lab, K)
        X = v.reshape(-1, 1)
        want = brute.best_split(X, lab, list(range(n)), [0])
        if want is None:
            assert got is None
        else:
            assert got is not None and got[0] == want[2], (trial, got, want)
            p, num, den = got
            # the oracle's integer score is 1 - Gini * n / ... : check Q = n (1 - G)
            from fractions import Fraction
            assert Fraction(num, den) == n * (1 - want[0])


# (1.0, 1.0000100135803223) is a candidate pair in double precision
# ((double)1 + 1e-5 < 1.00001001358) but not when the gap test is done in
# float, where 1 + 1e-5f rounds up to 1.00001001358.  No float pair can make
# (double)a + 1e-5 == (double)b exactly (1e-5's 53-bit mantissa has no run of
# 29 zero bits), so "<" and "<=" in the gap test cannot be told apart.
GAP_A = np.float32(1.0)
GAP_B = np.float32(1.0000100135803223)


def test_candidate_gap_is_double_precision(orc):
    assert float(GAP_A) + 1e-5 < float(GAP_B)
    v = np.array([GAP_A, GAP_B], np.float32)
    assert orc.scan_feature(v, np.array([0, 1]), 2) == (0, 2, 1)
    below = np.array([GAP_A, np.float32(1.000009)], np.float32)
    assert orc.scan_feature(below, np.array([0, 1]), 2) is None


def gap_case():
    X = np.array([[GAP_A], [GAP_B], [GAP_B], [GAP_A]], np.float32)
    return {"X": X, "y": np.array([0, 1, 1, 0]), "K": 2, "m": 1, "mss": 1, "max_depth": 0,
            "identity": True, "seed": 0, "boot": False, "n_trees": 1}


TINY_SEEDS = list(range(120))


def test_fit_matches_bruteforce_tiny(orc):
    for s in TINY_SEEDS:
        case = datasets.tiny_case(s)
        got = helpers.oracle_forest(case, orc)
        want = helpers.brute_forest(case, orc)
        d = helpers.forest_diff(got, want)
        assert d is None, "seed %d: %s" % (s, d)


def test_fit_gap_pair_dataset(orc):
    case = gap_case()
    got = helpers.oracle_forest(case, orc)
    assert len(got["feature"]) == 3          # the double-precision gap admits the split
    assert helpers.forest_diff(got, helpers.brute_forest(case, orc)) is None


def test_predict_matches_bruteforce(orc):
    for s in range(40):
        case = datasets.tiny_case(1000 + s)
        forest = helpers.oracle_forest(case, orc)
        rng = np.random.default_rng(s)
        Xq = np.concatenate([case["X"], rng.normal(1, 2, size=(30, case["X"].shape[1]))]).astype(np.float32)
        got, votes = orc.predict(forest, Xq, case["K"], return_votes=True)
        assert np.array_equal(got, brute.predict(forest, Xq, case["K"]))
        assert np.all(votes.sum(axis=1) == case["n_trees"])


@pytest.mark.parametrize("name", ["digits", "iris", "diabetes"])
def test_memorize(orc, name):
    """The reference's memorize tests (test/test_digits.py:7-17, test_iris.py:7-14,
    test_diabetes_classify.py:7-14): bootstrap=False forests reproduce every
    training label."""
    X, y = datasets.load_sklearn(name)
    u, yc = np.unique(y, return_inverse=True)
    m = int(np.ceil(np.sqrt(X.shape[1])))
    forest = orc.fit(X, yc, len(u), 6, m, seed=7)
    assert np.array_equal(orc.predict(forest, X, len(u)), yc)


@pytest.mark.parametrize("name", ["digits", "iris", "diabetes"])
def test_accuracy_vs_sklearn(orc, name):
    """test/helpers.py:4-28: train on the first half, score the second half,
    total over repeats within 2% of scikit-learn's forest."""
    from sklearn.ensemble import RandomForestClassifier

    X, y = datasets.load_sklearn(name)
    u, yc = np.unique(y, return_inverse=True)
    n = len(yc) // 2
    m = int(np.ceil(np.sqrt(X.shape[1])))
    ours = theirs = 0.0
    for rep in range(5):
        ids = datasets.bootstrap_ids(n, 11, 100 + rep)
        forest = orc.fit(X[:n], yc[:n], len(u), 11, m, boot_ids=ids, seed=rep)
        ours += np.mean(orc.predict(forest, X[n:], len(u)) == yc[n:])
        sk = RandomForestClassifier(n_estimators=11, random_state=rep).fit(X[:n], yc[:n])
        theirs += sk.score(X[n:], yc[n:])
    assert ours >= 0.98 * theirs, (ours / 5, theirs / 5)


MUTANTS = {
    1: "left sum of squares grows by 2c instead of 2c+1",
    2: "wrong sign in the right sum of squares",
    3: "candidate gap tested in float instead of double",
    4: "threshold is the left value instead of the midpoint",
    5: "ties go to the highest feature id",
    6: "right side of the partition reversed (unstable)",
    7: "leaf label from the last sample instead of the first",
    8: "left count off by one",
    9: "constant features not skipped",
}


@pytest.fixture(scope="module")
def mutant_libs():
    oracle.build(mutants=True)
    d = os.path.join(os.path.dirname(oracle.__file__), "mutants")
    return {k: oracle.Oracle(os.path.join(d, "liboracle_m%d.so" % k)) for k in MUTANTS}


def _caught(mut, orc):
    cases = [gap_case()] + [datasets.tiny_case(s) for s in TINY_SEEDS]
    for case in cases:
        if case["mss"] > 1:
            case = dict(case, mss=5)
        got = helpers.oracle_forest(case, mut)
        want = helpers.brute_forest(case, orc)
        if helpers.forest_diff(got, want) is not None:
            return True
        Xq = case["X"]
        if not np.array_equal(mut.predict(got, Xq, case["K"]), brute.predict(want, Xq, case["K"])):
            return True
    return False


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_planted_mistake_is_caught(orc, mutant_libs, k):
    assert _caught(mutant_libs[k], orc), "planted mistake %d (%s) not caught" % (k, MUTANTS[k])


def test_threaded_fit_equals_sequential():
    """Trees grown by several host threads at once form the same forest, in
    tree order, as trees grown one after another."""
    import oracle
    from cudatree_amd import datasets

    X, y, _, _ = datasets.covtype_like(seed=4, n_train=3000, n_test=1)
    orc = oracle.get()
    ids = orc.bootstrap_forest_ids(11, 9, len(y))
    a = orc.fit(X, y, 7, 9, 7, boot_ids=ids, seed=11)
    for threads in (2, 4, 16):
        b = orc.fit(X, y, 7, 9, 7, boot_ids=ids, seed=11, threads=threads)
        for k in a:
            assert np.array_equal(a[k], b[k]), (threads, k)
