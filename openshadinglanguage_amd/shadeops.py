"""Batched shadeop calls on torch device tensors (SoA planes).

Mirrors the reference's wide shadeop entry points
(src/liboslexec/wide/wide_opnoise_impl_xmacro.h:57-127): one call shades n
points; inputs are [planes, n] float32 tensors on the GPU.
"""
import ctypes

import torch

from ._capi import check, lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _planes(t):
    if not (t.is_cuda and t.dtype == torch.float32 and t.dim() == 2 and t.is_contiguous()):
        raise ValueError("expected a contiguous float32 [planes, n] CUDA tensor")
    return t


def noise(name, outdim, inp, derivs=False):
    """noise(name, P) over n points. inp: [indim (x3 with derivs), n]."""
    _planes(inp)
    planes, n = inp.shape
    indim = planes // 3 if derivs else planes
    if derivs and planes % 3:
        raise ValueError("Dual2 input needs 3*indim planes")
    out = torch.empty(((3 if derivs else 1) * outdim, n), device=inp.device, dtype=torch.float32)
    check(lib().oslmi_noise(name.encode(), outdim, indim, int(derivs), _ptr(inp), _ptr(out), n,
                            _stream()))
    return out


def pnoise(name, outdim, inp, period, derivs=False):
    """pnoise(name, P, period) over n points. inp as for noise(); period:
    [indim, n] float32 (one period per input dimension and point)."""
    _planes(inp)
    _planes(period)
    planes, n = inp.shape
    indim = planes // 3 if derivs else planes
    if derivs and planes % 3:
        raise ValueError("Dual2 input needs 3*indim planes")
    if tuple(period.shape) != (indim, n):
        raise ValueError("period needs [indim, n] planes")
    out = torch.empty(((3 if derivs else 1) * outdim, n), device=inp.device, dtype=torch.float32)
    check(lib().oslmi_pnoise(name.encode(), outdim, indim, int(derivs), _ptr(inp), _ptr(period),
                             _ptr(out), n, _stream()))
    return out


def hash_op(inp):
    _planes(inp)
    out = torch.empty(inp.shape[1], device=inp.device, dtype=torch.int32)
    check(lib().oslmi_hash(inp.shape[0], _ptr(inp), _ptr(out), inp.shape[1], _stream()))
    return out
