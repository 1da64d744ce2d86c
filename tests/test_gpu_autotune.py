"""Autotuner: every candidate schedule grows the oracle's trees; the tuner
returns the fastest one."""
import pytest

import helpers
import oracle
from cudatree_amd import _lib, autotune, datasets

pytestmark = pytest.mark.gpu


def test_every_setting_grows_the_oracle_trees():
    """Every setting of the default grid (scan tile width, reshuffle task
    size, bitmap staging, class-count layout) and the hybrid schedule at
    every bfs_threshold grow the oracle's forest bit for bit."""
    X, y, _, _ = datasets.covtype_like(seed=31, n_train=30000, n_test=1)
    ctx = _lib.default_context(0)
    ds = _lib.Dataset(ctx, X, y, 7)
    orc = oracle.get()
    want = orc.fit(X, y, 7, 3, 7, boot_ids=orc.bootstrap_forest_ids(77, 3, len(y)), seed=77)
    settings = autotune.candidate_settings(K=7)
    assert any(s.get("hist_layout") == _lib.CTA_HIST_LDS for s in settings)
    for setting in settings:
        o = _lib.default_opts()
        o.seed = 77
        o.max_features = 7
        o.bootstrap = 1
        autotune.apply(o, setting)
        print("setting %s" % setting, flush=True)
        got = _lib.Forest.fit(ctx, ds, 3, o, None).export()
        assert helpers.forest_diff(got, want) is None, setting


def test_many_classes_drop_register_layout():
    assert all(s.get("hist_layout", _lib.CTA_HIST_LDS) == _lib.CTA_HIST_LDS
               for s in autotune.candidate_settings(K=100))


def test_tune_picks_fastest_and_caches():
    X, y, _, _ = datasets.covtype_like(seed=32, n_train=20000, n_test=1)
    best, table = autotune.tune(X, y, 7, 7, probe_trees=2, repeats=1, hybrid=False,
                                grid={"scan_chunk": [2048, 8192], "resh_chunk": [2048], "resh_global": [0],
                                      "hist_layout": [1, 2]})
    assert len(table) == 4
    assert best == table[0][0] and all(table[i][1] <= table[i + 1][1] for i in range(len(table) - 1))
    assert autotune.tune(X, y, 7, 7)[0] == best   # cached per problem shape


def test_facade_autotune_same_forest():
    import cudatree_amd

    X, y, _, _ = datasets.covtype_like(seed=33, n_train=15000, n_test=1)
    a = cudatree_amd.RandomForestClassifier(n_estimators=3, random_state=5, autotune=True).fit(X, y)
    b = cudatree_amd.RandomForestClassifier(n_estimators=3, random_state=5).fit(X, y)
    assert a.tuned_ is not None
    assert helpers.forest_diff(a.export_trees(), b.export_trees()) is None
