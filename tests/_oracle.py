"""ctypes binding of the CPU oracle (oracle/build/liboslref.so).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
this; the product package never does.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "build", "liboslref.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        L.oref_inthash.argtypes = [P, ctypes.c_int]
        L.oref_inthash.restype = ctypes.c_uint32
        L.oref_bjfinal.argtypes = [ctypes.c_uint32] * 3
        L.oref_bjfinal.restype = ctypes.c_uint32
        L.oref_noise.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, P, P, ctypes.c_int64]
        L.oref_noise.restype = ctypes.c_int
        L.oref_hash_op.argtypes = [ctypes.c_int, P, P, ctypes.c_int64]
        L.oref_hash_op.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def inthash(words):
    k = np.ascontiguousarray(np.asarray(words, dtype=np.uint32))
    return int(lib().oref_inthash(_p(k), len(k)))


def noise(name, outdim, inp, derivs=False):
    """inp: float32 array [planes, n] (planes = indim, or 3*indim with
    derivs: val, dx, dy blocks). Returns [outdim or 3*outdim, n]."""
    inp = np.ascontiguousarray(inp, dtype=np.float32)
    planes, n = inp.shape
    indim = planes // 3 if derivs else planes
    out = np.zeros(((3 if derivs else 1) * outdim, n), dtype=np.float32)
    rc = lib().oref_noise(name.encode(), outdim, indim, int(derivs), _p(inp),
                          _p(out), n)
    if rc != 0:
        raise ValueError("oracle rejected noise request %r" % name)
    return out


def hash_op(inp):
    inp = np.ascontiguousarray(inp, dtype=np.float32)
    out = np.zeros(inp.shape[1], dtype=np.int32)
    assert lib().oref_hash_op(inp.shape[0], _p(inp), _p(out), inp.shape[1]) == 0
    return out
