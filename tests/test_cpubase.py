"""The CPU baseline's throughput builds of the oracle (oracle/cpubase,
timed by bench.py, never a checker) load and export the entry points the
benchmark calls (a missing symbol would only show on the GPU host)."""
import ctypes
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "..", "oracle", "build")


@pytest.mark.parametrize("level", ["v3", "v4"])
def test_w16_libraries_load(level):
    lib = ctypes.CDLL(os.path.join(BUILD, "liboslref_w16_%s.so" % level))
    for sym in ("oref_w16_shade_c1", "oref_w16_shade_c2"):
        assert hasattr(lib, sym)
