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


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of the ABI structs have the header's size and
    field offsets (compiled with the host C compiler, no GPU)."""
    from cudatree_amd import _lib

    structs = {"cta_fit_opts": _lib.FitOpts, "cta_server_opts": _lib.ServerOpts,
               "cta_server_stats": _lib.ServerStats, "cta_kernel_stat": _lib.KernelStat}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "cudatree_amd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append('  printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for field, _ in py._fields_:
            lines.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, field, cname, field))
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                         text=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert int(got["%s.%s" % (cname, field)]) == getattr(py, field).offset, (cname, field)


def test_fit_estimate_host_only():
    """cta_fit_estimate is a host computation: it grows with rows, trees and
    features, a batch cap bounds the resident trees, bad sizes are refused."""
    from cudatree_amd import _lib

    o = _lib.default_opts()
    o.max_features = 7
    ds1, ws1 = _lib.fit_estimate(20000, 54, 7, 10, o)
    ds2, ws2 = _lib.fit_estimate(40000, 54, 7, 10, o)
    _, ws3 = _lib.fit_estimate(20000, 54, 7, 20, o)
    assert 0 < ds1 < ds2 and 0 < ws1 < ws2 and ws1 < ws3
    # the dataset alone holds X_T, the presorted rows and the copies of X and y
    assert ds1 >= 20000 * 54 * (4 + 4 + 4)
    o.trees_per_batch = 2
    _, ws4 = _lib.fit_estimate(20000, 54, 7, 20, o)
    assert ws4 < ws3
    with pytest.raises(_lib.CtaError):
        _lib.fit_estimate(0, 54, 7, 10, o)
    with pytest.raises(_lib.CtaError):
        _lib.fit_estimate(100, 54, 7, 0, o)
