"""The C ABI library loads without a GPU and exports every function that
include/glimp.h declares; the Python binding covers exactly that set."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "glimp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set(re.findall(r"\b(glimp_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_parses():
    names = _header_functions()
    assert "glimp_draw_meshlets" in names and "glimp_batch_launch" in names
    assert len(names) >= 25


def test_library_exports_header_symbols():
    import glimpsw_amd as g

    path = g.library_path()
    if not os.path.exists(path):
        pytest.skip("libglimp.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    missing = [n for n in sorted(_header_functions()) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_covers_header():
    import glimpsw_amd as g

    assert set(g._SIGNATURES) == _header_functions()
