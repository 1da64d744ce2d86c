"""Versioned forest images and façade pickling, on the host (no GPU): the
byte image of cta_forest_save as written by the Python codec, round trips,
and the rejection of other versions and malformed images.  The device side
(cta_forest_save / cta_forest_load) is checked against this codec in
tests/test_gpu_serialize.py."""
import pickle

import numpy as np
import pytest

import oracle
from cudatree_amd import RandomForestClassifier, _lib, datasets


@pytest.fixture(scope="module")
def forest():
    X, y = datasets.digits_like(seed=5, n=400)
    orc = oracle.get()
    arrays = orc.fit(X, y, 10, 3, 8, boot_ids=orc.bootstrap_forest_ids(9, 3, len(y)), seed=9)
    return X, y, arrays


def test_image_layout_and_round_trip(forest):
    _, _, a = forest
    img = _lib.forest_to_bytes(a, 64, 10)
    n, T = len(a["feature"]), len(a["tree_off"]) - 1
    assert len(img) == 48 + 8 * (T + 1) + 28 * n
    assert img[:8] == b"CTAFORST" and int.from_bytes(img[8:12], "little") == _lib.FOREST_FORMAT
    assert int.from_bytes(img[12:16], "little") == 48
    back, F, K = _lib.forest_from_bytes(img)
    assert (F, K) == (64, 10)
    for k in ("tree_off",) + _lib.FOREST_FIELDS:
        assert np.array_equal(back[k], np.asarray(a[k])), k
    assert back["threshold"].view(np.uint32).tolist() == np.asarray(a["threshold"]).view(np.uint32).tolist()


def test_image_rejects_other_versions_and_damage(forest):
    _, _, a = forest
    img = bytearray(_lib.forest_to_bytes(a, 64, 10))
    bad = bytearray(img)
    bad[8:12] = (2).to_bytes(4, "little")
    with pytest.raises(ValueError, match="version 2"):
        _lib.forest_from_bytes(bytes(bad))
    bad = bytearray(img)
    bad[0:8] = b"NOTAFRST"
    with pytest.raises(ValueError, match="magic"):
        _lib.forest_from_bytes(bytes(bad))
    with pytest.raises(ValueError):
        _lib.forest_from_bytes(bytes(img[:-4]))
    with pytest.raises(ValueError):
        _lib.forest_from_bytes(bytes(img[:20]))


def _fitted_facade(forest):
    X, y, a = forest
    rf = RandomForestClassifier(n_estimators=3, max_features=8, random_state=9)
    rf.compt_table = np.array(["c%d" % k for k in range(10)])
    rf.n_labels, rf.n_features, rf.max_features_ = 10, 64, 8
    rf._arrays = {k: np.asarray(v) for k, v in a.items()}
    return rf


def test_pickle_carries_version_labels_and_image(forest):
    rf = _fitted_facade(forest)
    state = rf.__getstate__()
    assert state["state_version"] == RandomForestClassifier.STATE_VERSION
    assert isinstance(state["forest_image"], bytes) and "_arrays" not in state
    g = pickle.loads(pickle.dumps(rf))
    assert list(g.compt_table) == list(rf.compt_table)
    got, want = g.export_trees(), rf.export_trees()
    for k in ("tree_off",) + _lib.FOREST_FIELDS:
        assert np.array_equal(got[k], want[k]), k


def test_pickle_rejects_unknown_state_version(forest):
    rf = _fitted_facade(forest)
    state = rf.__getstate__()
    for version in (None, 0, RandomForestClassifier.STATE_VERSION + 1):
        st = dict(state)
        if version is None:
            del st["state_version"]
        else:
            st["state_version"] = version
        g = RandomForestClassifier.__new__(RandomForestClassifier)
        with pytest.raises(ValueError, match="state version"):
            g.__setstate__(st)
