"""Host-side parts of the layered closure (no GPU): the library's baked
energy tables equal the oracle's, and invalid requests are rejected before
any device work."""
import ctypes

import numpy as np

import _oracle as O
from openshadinglanguage_amd import OSLMIError, check, lib
from openshadinglanguage_amd import workloads as W


def test_library_tables_equal_oracle_tables():
    out = np.zeros(8448, np.float32)
    check(lib().oslmi_bsdl_tables(out.ctypes.data_as(ctypes.c_void_p), 8448))
    ref = O.mtx_tables()
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_tables_buffer_size_checked():
    out = np.zeros(10, np.float32)
    assert lib().oslmi_bsdl_tables(out.ctypes.data_as(ctypes.c_void_p), 10) != 0


def test_transmission_rejected():
    m = W.layered_material()
    m.coat_refr_tint[:] = [0.5, 0.5, 0.5]
    rc = lib().oslmi_layered_closure(ctypes.byref(m), 16, None, None, None, None, None, None,
                                     None, None)
    assert rc == -3  # OSLMI_ERR_UNSUPPORTED
    assert b"transmission" in lib().oslmi_last_error()


def test_null_buffers_rejected():
    m = W.layered_material()
    rc = lib().oslmi_layered_closure(ctypes.byref(m), 16, None, None, None, None, None, None,
                                     None, None)
    assert rc == -1
    rc = lib().oslmi_layered_eval_local(ctypes.byref(m), 16, None, None, None)
    assert rc == -1
    assert lib().oslmi_layered_eval_local(ctypes.byref(m), 0, None, None, None) == 0


def test_queries_shape_and_ranges():
    q = W.surrogate_queries(4096, seed=6)
    assert q.shape == (14, 4096) and q.dtype == np.float32
    assert np.allclose((q[:3].astype(np.float64) ** 2).sum(0), 1, atol=1e-6)
    assert np.allclose(q[3] ** 2 + q[4] ** 2, 1, atol=1e-6) and q[4].min() >= 0
    assert q[5].min() >= 1.3 and q[5].max() <= 1.8
    try:
        check(-1)
    except OSLMIError:
        pass
