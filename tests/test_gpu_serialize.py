"""Device side of the forest image (needs an MI355X): cta_forest_save writes
the bytes of the host codec, cta_forest_load restores a forest that
predicts as the oracle, and other versions are refused."""
import pickle

import numpy as np
import pytest

import oracle
from cudatree_amd import RandomForestClassifier, _lib, datasets

pytestmark = pytest.mark.gpu


def test_save_load_matches_codec_and_oracle():
    ctx = _lib.default_context(0)
    orc = oracle.get()
    X, y, Xt, _ = datasets.covtype_like(seed=51, n_train=12000, n_test=4000)
    ds = _lib.Dataset(ctx, X, y, 7)
    o = _lib.default_opts()
    o.seed, o.max_features, o.bootstrap = 53, 7, 1
    print("save/load fit", flush=True)
    f = _lib.Forest.fit(ctx, ds, 6, o, None)
    arrays = f.export()
    img = f.save()
    assert img == _lib.forest_to_bytes(arrays, 54, 7)
    g = _lib.Forest.load(ctx, img)
    assert (g.n_trees, g.n_nodes, g.F, g.K) == (6, len(arrays["feature"]), 54, 7)
    back = g.export()
    for k in ("tree_off",) + _lib.FOREST_FIELDS:
        assert np.array_equal(back[k], arrays[k]), k
    want = orc.predict(arrays, Xt, 7, return_votes=True)
    got = g.predict(Xt, votes=True)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    bad = bytearray(img)
    bad[8:12] = (_lib.FOREST_FORMAT + 1).to_bytes(4, "little")
    with pytest.raises(_lib.CtaError) as e:
        _lib.Forest.load(ctx, bytes(bad))
    assert e.value.code == 4  # CTA_EUNSUPPORTED
    with pytest.raises(_lib.CtaError):
        _lib.Forest.load(ctx, img[:-8])


def test_facade_pickle_predicts_after_restore():
    x, y = datasets.digits_like(seed=3, n=900)
    labels = np.array(["k%02d" % v for v in y])
    f = RandomForestClassifier(n_estimators=5, random_state=12).fit(x, labels)
    print("facade pickle", flush=True)
    g = pickle.loads(pickle.dumps(f))
    assert np.array_equal(g.predict(x), f.predict(x))
    assert np.array_equal(g.predict_proba(x[:50]), f.predict_proba(x[:50]))
