"""Serving loop on the GPU: each result is the same alone and in the mix,
admission by HBM accounting, coalesced and unbatched launches."""
import threading

import numpy as np
import pytest

import helpers
import oracle
from cudatree_amd import _lib, datasets

pytestmark = pytest.mark.gpu


def fit_opts(seed, m=7, max_depth=0):
    o = _lib.default_opts()
    o.seed = seed
    o.max_features = m
    o.max_depth = max_depth
    o.bootstrap = 1
    return o


@pytest.fixture(scope="module")
def data():
    X, y, Xt, yt = datasets.covtype_like(seed=21, n_train=20000, n_test=20000)
    return X, y, Xt


@pytest.fixture(scope="module")
def base_forest(data):
    X, y, _ = data
    ctx = _lib.default_context(0)
    ds = _lib.Dataset(ctx, X, y, 7)
    return _lib.Forest.fit(ctx, ds, 20, fit_opts(3), None)


def _requests(Xt, n, seed, largest=3000):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        r = int(np.exp(rng.uniform(0, np.log(largest))))
        a = int(rng.integers(0, len(Xt) - r))
        out.append(Xt[a:a + r])
    return out


# (coalesce, split_rows): one lane per request; coalesced with the short and
# long lanes split at 512 rows; coalesced on one lane
@pytest.mark.parametrize("coalesce,split", [(True, 512), (False, 0), (True, -1)])
def test_predict_alone_vs_in_mix(data, base_forest, coalesce, split):
    X, y, Xt = data
    reqs = _requests(Xt, 60, 5, largest=12000)
    alone = [base_forest.predict(r) for r in reqs]
    # the served labels are the oracle's traversal and vote of the same trees
    trees = base_forest.export()
    orc = oracle.get()
    for r, a in zip(reqs, alone):
        assert np.array_equal(a, orc.predict(trees, r, 7))
    fit_jobs = [(X[:5000 * (k + 1)], y[:5000 * (k + 1)], 10 + 5 * k, 100 + k) for k in range(3)]
    fit_alone = []
    ctx = _lib.default_context(0)
    for Xf, yf, T, seed in fit_jobs:
        ds = _lib.Dataset(ctx, Xf, yf, 7)
        fit_alone.append(_lib.Forest.fit(ctx, ds, T, fit_opts(seed), None).export())
    srv = _lib.Server(coalesce=coalesce, fit_workers=1, split_rows=split)
    try:
        fid = srv.add_forest(base_forest)
        tickets = [None] * len(reqs)
        fit_tickets = []

        def client(part):
            for i in range(part, len(reqs), 4):
                tickets[i] = srv.submit_predict(fid, reqs[i])

        print("serving mix coalesce=%d split_rows=%d" % (coalesce, split), flush=True)
        for Xf, yf, T, seed in fit_jobs:
            fit_tickets.append(srv.submit_fit(Xf, yf, 7, T, fit_opts(seed)))
        threads = [threading.Thread(target=client, args=(p,)) for p in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for i, t in enumerate(tickets):
            st, labels = srv.wait(t)
            assert st == 0
            assert np.array_equal(labels, alone[i]), "request %d differs in the mix" % i
        for k, t in enumerate(fit_tickets):
            st, new_fid = srv.wait(t)
            assert st == 0 and new_fid >= 0
            assert helpers.forest_diff(srv.export_forest(new_fid), fit_alone[k]) is None
        s = srv.stats()
        assert s["predict_requests"] == len(reqs) and s["fit_jobs"] == 3
        if coalesce:
            assert s["predict_launches"] <= s["predict_requests"]
        else:
            assert s["predict_launches"] == s["predict_requests"]
        assert s["p99_ms"] >= s["p50_ms"] > 0 and s["hbm_peak"] <= s["hbm_budget"]
    finally:
        srv.close()


@pytest.mark.parametrize("coalesce", [True, False])
def test_submit_many_matches_single(data, base_forest, coalesce):
    """Requests queued in one call give the labels of single predicts (and
    the oracle's), across both lanes."""
    _, _, Xt = data
    reqs = _requests(Xt, 300, 9, largest=9000)
    trees = base_forest.export()
    orc = oracle.get()
    srv = _lib.Server(coalesce=coalesce, split_rows=256)
    try:
        fid = srv.add_forest(base_forest)
        print("submit_many coalesce=%d" % coalesce, flush=True)
        tickets = srv.submit_many(fid, reqs[:150]) + srv.submit_many(fid, reqs[150:])
        for r, t in zip(reqs, tickets):
            st, labels = srv.wait(t)
            assert st == 0 and np.array_equal(labels, orc.predict(trees, r, 7))
        s = srv.stats()
        assert s["predict_requests"] == len(reqs)
        assert s["predict_launches"] < len(reqs) if coalesce else s["predict_launches"] == len(reqs)
        with pytest.raises(_lib.CtaError):
            srv.submit_many(fid, [reqs[0], np.zeros((0, 54), np.float32)])
    finally:
        srv.close()


@pytest.mark.parametrize("coalesce", [True, False])
def test_submit_rows_and_wait_many(data, base_forest, coalesce):
    """Requests laid back to back in one array, queued in one call and
    waited for in one call: every request's labels equal the oracle's, at
    the offsets of its rows; unknown tickets are refused and a zero timeout
    reports pending tickets as -1 without forgetting them."""
    _, _, Xt = data
    reqs = _requests(Xt, 200, 21, largest=6000)
    R = np.array([len(r) for r in reqs], np.int64)
    X = np.concatenate(reqs)
    trees = base_forest.export()
    want = oracle.get().predict(trees, X, 7)
    srv = _lib.Server(coalesce=coalesce, split_rows=512)
    try:
        fid = srv.add_forest(base_forest)
        print("submit_rows coalesce=%d" % coalesce, flush=True)
        t1, lab1 = srv.submit_rows(fid, X[:R[:120].sum()], R[:120])
        t2, lab2 = srv.submit_rows(fid, X[R[:120].sum():], R[120:])
        tickets = np.concatenate([t1, t2])
        early = srv.wait_many(tickets, timeout_us=0)
        assert set(early.tolist()) <= {-1, 0}
        st = srv.wait_many(tickets[early < 0])
        assert (st == 0).all()
        assert np.array_equal(np.concatenate([lab1, lab2]), want)
        s = srv.stats()
        assert s["predict_requests"] == len(reqs)
        assert s["predict_launches"] < len(reqs) if coalesce else s["predict_launches"] == len(reqs)
        with pytest.raises(_lib.CtaError):
            srv.wait_many(tickets[:1])  # reported already: forgotten
        with pytest.raises(ValueError):
            srv.submit_rows(fid, X[:10], np.array([4, 5]))
    finally:
        srv.close()


def test_fit_admission_by_hbm_budget(data):
    X, y, _ = data
    srv = _lib.Server(hbm_budget=256 << 20)
    try:
        t = srv.submit_fit(X, y, 7, 50, fit_opts(1))       # needs far more than 256 MB
        st, _ = srv.wait(t)
        assert st == 2  # CTA_ENOMEM: rejected, never started
        t = srv.submit_fit(X[:2000], y[:2000], 7, 5, fit_opts(1))   # small job fits
        st, fid = srv.wait(t)
        assert st == 0 and fid >= 0
        s = srv.stats()
        assert s["fit_rejected"] == 1 and s["fit_jobs"] == 1 and s["hbm_peak"] <= 256 << 20
        # the device's free memory fell by no more than the job was admitted
        # with (a small job may be served from memory the runtime keeps)
        assert s["fit_mem_ratio_max"] <= 1.0, s
    finally:
        srv.close()


def test_jobs_queue_until_memory_frees(data):
    X, y, _ = data
    import ctypes

    # budget for about one of these jobs at a time (the library's estimate)
    one = sum(_lib.fit_estimate(len(y), X.shape[1], 7, 8, fit_opts(1)))
    srv = _lib.Server(hbm_budget=int(one * 1.5), fit_workers=2)
    try:
        ts = [srv.submit_fit(X, y, 7, 8, fit_opts(s)) for s in (1, 2, 3)]
        for t in ts:
            st, fid = srv.wait(t)
            assert st == 0
        s = srv.stats()
        assert s["fit_jobs"] == 3 and s["hbm_peak"] <= int(one * 1.5)
        assert s["fit_mem_ratio_max"] <= 1.0, s
    finally:
        srv.close()


def test_bad_requests_are_rejected(base_forest):
    srv = _lib.Server()
    try:
        fid = srv.add_forest(base_forest)
        with pytest.raises(_lib.CtaError):
            srv.submit_predict(fid + 7, np.zeros((3, 54), np.float32))
        with pytest.raises(_lib.CtaError):
            srv.submit_predict(fid, np.zeros((3, 5), np.float32))
        srv.drop_forest(fid)
        with pytest.raises(_lib.CtaError):
            srv.submit_predict(fid, np.zeros((3, 54), np.float32))
    finally:
        srv.close()
