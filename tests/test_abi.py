"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h declares (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cudatree_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cta_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from cudatree_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        subprocess.run(["make", "-C", os.path.join(ROOT, "cudatree_amd", "csrc")], check=True)
    return ctypes.CDLL(_lib.LIB_PATH)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "cta_fit" in names and "cta_predict" in names and len(names) >= 20


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from cudatree_amd import _lib

    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_error_paths_without_gpu():
    """Argument checks fail cleanly (no device work): NULL context."""
    from cudatree_amd import _lib

    lib = _lib.load()
    assert lib.cta_version() >= 100
    rc = lib.cta_ctx_synchronize(None)
    assert rc == 1 and b"NULL" in lib.cta_last_error()
