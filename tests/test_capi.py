"""The C-ABI library loads and exports every symbol include/oslmi.h declares
(no compute calls: these run without a GPU)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"^\s*(?:const\s+)?[\w\*\s]+?\b(oslmi_\w+)\s*\(", src, re.M)
    return sorted(set(names))


def test_library_exports_declared_symbols():
    import ctypes
    from openshadinglanguage_amd._capi import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.skip("liboslmi.so not built")
    L = ctypes.CDLL(LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert declared(), "no declarations found"
    assert not missing, missing


def test_version_string():
    from openshadinglanguage_amd import lib
    assert b"gfx950" in lib().oslmi_version()
