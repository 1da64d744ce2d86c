"""Class-value keys for the device label compaction (host side, no GPU):
key order is class order, equal classes share a key, and the classes decode
back to the input dtype -- the contract of np.unique(return_inverse=True)
(random_forest.py:85-99) that cta_dataset_create_keys computes on the
device."""
import numpy as np
import pytest

from cudatree_amd import _lib


def _unique_via_keys(target):
    keys, bits, decode = _lib.class_keys(target)
    assert keys.dtype == (np.uint32 if bits == 32 else np.uint64)
    uk, inv = np.unique(keys, return_inverse=True)
    return decode(uk), inv


CASES = {
    "uint8": np.array([3, 0, 255, 3, 7], np.uint8),
    "bool": np.array([True, False, True]),
    "int8": np.array([-128, 127, 0, -1, -128], np.int8),
    "int32": np.array([-(2**31), 2**31 - 1, 0, -5, 17, -5], np.int32),
    "uint32": np.array([2**32 - 1, 0, 5], np.uint32),
    "int64": np.array([-(2**63), 2**63 - 1, 0, -1, 2**40], np.int64),
    "uint64": np.array([2**64 - 1, 0, 2**63, 5], np.uint64),
    "float32": np.array([1.5, -0.0, 0.0, -3.25, np.inf, -np.inf, 1e-30, -1e-30], np.float32),
    "float64": np.array([2.5, -7.0, 1e300, -1e-300, 0.0, 2.5], np.float64),
    "float16": np.array([1.0, -2.0, 0.5], np.float16),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_keys_follow_np_unique(name):
    t = CASES[name]
    classes, inv = _unique_via_keys(t)
    want_c, want_inv = np.unique(t, return_inverse=True)
    assert classes.dtype == t.dtype
    assert np.array_equal(classes, want_c)
    assert np.array_equal(inv, want_inv)


def test_random_labels_match_np_unique():
    rng = np.random.default_rng(3)
    for dt in (np.int16, np.int64, np.float32, np.float64):
        t = (rng.integers(-50, 50, size=5000) * 3).astype(dt)
        classes, inv = _unique_via_keys(t)
        want_c, want_inv = np.unique(t, return_inverse=True)
        assert np.array_equal(classes, want_c) and np.array_equal(inv, want_inv)


def test_signed_zero_is_one_class_and_nan_is_refused():
    keys, _, _ = _lib.class_keys(np.array([-0.0, 0.0], np.float64))
    assert keys[0] == keys[1]
    with pytest.raises(ValueError):
        _lib.class_keys(np.array([1.0, np.nan]))


def test_strings_have_no_keys():
    assert _lib.class_keys(np.array(["a", "b"])) is None
    assert _lib.class_keys(np.array([object(), object()], dtype=object)) is None
