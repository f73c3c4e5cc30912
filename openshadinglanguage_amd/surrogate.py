"""Learned surrogate of the layered closure (host side).

The network, its training kernels and the fused inference kernel live in
the HIP library (csrc/oslmi_surrogate.hip, C ABI oslmi_surrogate_*). This
module drives them:

  * training data: queries from the closure workload's distribution with L
    in the upper hemisphere (the network's domain), labelled on the GPU by
    the exact closure kernel (oslmi_layered_eval_local);
  * training: minibatch Adam with a cosine learning-rate schedule, every
    step one oslmi_surrogate_train_step (BF16 MFMA forward + backward);
  * calibration: the network's relative error per routing cell on held-out
    queries; cells above the error budget go to the exact closure;
  * inference: oslmi_surrogate_infer (network + exact fallback).
"""
import ctypes
import math
import time

import numpy as np

from . import workloads as W
from ._capi import check, lib

NIN, NOUT = 14, 4
# (x - offset) * scale maps every input into about [-1, 1]
IN_OFFSET = [0.0, 0.0, 0.0, 0.0, 0.0, 1.55, 0.525, 0.55, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5]
IN_SCALE = [1.0, 1.0, 1.0, 1.0, 1.0, 4.0, 2.0 / 0.95, 2.0 / 0.9, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0]
# training domain of the 9 material parameters (the closure workload's ranges)
DOMAIN_LO = [1.3, 0.05, 0.1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
DOMAIN_HI = [1.8, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0]
SHAPES = [("W1", (64, 32)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)),
          ("b3", (64,)), ("W4", (64, 64)), ("b4", (64,)), ("W5", (32, 64)), ("b5", (32,))]
NCELL = 64


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _f(vals):
    return (ctypes.c_float * len(vals))(*[float(v) for v in vals])


def init_params(seed=0):
    """Uniform He-style initialisation (fixed seed); zero padding rows and
    columns."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            fan_in = NIN if name == "W1" else 64
            lim = 0.5 * math.sqrt(6.0 / fan_in)
            w = rng.uniform(-lim, lim, size=shape)
            if name == "W1":
                w[:, NIN:] = 0.0
            if name == "W5":
                w[NOUT:, :] = 0.0
            parts.append(w.ravel())
        else:
            parts.append(np.zeros(shape).ravel())
    return np.concatenate(parts).astype(np.float32)


def upper_hemisphere_queries(n, seed=5):
    """Closure-workload queries restricted to L.z > 0 (the network's
    domain): L uniform on the upper hemisphere, the rest as
    workloads.surrogate_queries."""
    q = W.surrogate_queries(n, seed)
    q[2] = np.abs(q[2])
    q[2][q[2] == 0] = np.float32(1e-7)
    return q


class Surrogate:
    def __init__(self, max_batch=65536):
        self.L = lib()
        self.max_batch = max_batch
        self.h = self.L.oslmi_surrogate_create(max_batch)
        if not self.h:
            raise RuntimeError("oslmi_surrogate_create: %s" % self.L.oslmi_last_error().decode())
        self.out_mean = [0.0] * NOUT
        self.out_std = [1.0] * NOUT
        self._set_norm()
        self.set_params(init_params(0))
        self.step = 0

    def __del__(self):
        try:
            self.L.oslmi_surrogate_destroy(self.h)
        except Exception:
            pass

    def _set_norm(self):
        check(self.L.oslmi_surrogate_set_normalization(
            self.h, _f(IN_OFFSET), _f(IN_SCALE), _f(self.out_mean), _f(self.out_std),
            _f(DOMAIN_LO), _f(DOMAIN_HI)))

    def set_output_normalization(self, mean, std):
        self.out_mean = [float(v) for v in mean]
        self.out_std = [float(v) for v in std]
        self._set_norm()

    def set_params(self, params):
        params = np.ascontiguousarray(params, np.float32)
        check(self.L.oslmi_surrogate_set_params(self.h, params.ctypes.data_as(ctypes.c_void_p),
                                                params.size))
        self.step = 0

    def get_params(self):
        out = np.zeros(self.L.oslmi_surrogate_param_count(), np.float32)
        check(self.L.oslmi_surrogate_get_params(self.h, out.ctypes.data_as(ctypes.c_void_p),
                                                out.size))
        return out

    def grads(self, q, t):
        """Gradient (host array) and loss for one batch of device planes."""
        out = np.zeros(self.L.oslmi_surrogate_param_count(), np.float32)
        loss = ctypes.c_float()
        check(self.L.oslmi_surrogate_grads(self.h, q.shape[1], _p(q), _p(t),
                                           out.ctypes.data_as(ctypes.c_void_p),
                                           ctypes.byref(loss)))
        return out, float(loss.value)

    def fit_output_normalization(self, targets):
        """Centres the log targets (device tensor [4, n]) on their mean and
        keeps unit scale, so the squared loss weighs every output's relative
        error alike."""
        import torch
        y = torch.empty_like(targets)
        y[:3] = torch.log(torch.clamp(targets[:3], min=1e-6))
        y[3] = torch.log(torch.clamp(targets[3], min=1e-12))
        self.set_output_normalization(y.mean(1).tolist(), [1.0] * NOUT)

    def train(self, queries, targets, steps, batch=16384, lr=2e-3, lr_min=1e-5, seed=0,
              log_every=0, stream=None, progress=None):
        """Minibatch Adam over device planes queries [14, n], targets [4, n].
        progress(k) (optional) is called every 2000 steps."""
        import torch
        n = queries.shape[1]
        gen = torch.Generator(device=queries.device)
        gen.manual_seed(seed)
        qb = torch.empty((NIN, batch), device=queries.device)
        tb = torch.empty((NOUT, batch), device=queries.device)
        stream = stream or torch.cuda.current_stream()
        sp = ctypes.c_void_p(stream.cuda_stream)
        t0 = time.perf_counter()
        log = []
        for k in range(steps):
            idx = torch.randint(0, n, (batch,), device=queries.device, generator=gen)
            torch.index_select(queries, 1, idx, out=qb)
            torch.index_select(targets, 1, idx, out=tb)
            self.step += 1
            cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * k / max(1, steps)))
            check(self.L.oslmi_surrogate_train_step(self.h, batch, _p(qb), _p(tb), cur, 0.9,
                                                    0.999, 1e-8, self.step, sp))
            if log_every and (k % log_every == 0 or k == steps - 1):
                _, loss = self.grads(qb, tb)
                log.append((k, loss))
            if progress and k % 2000 == 0:
                progress(k)
        torch.cuda.synchronize()
        return {"steps": steps, "batch": batch, "seconds": time.perf_counter() - t0, "log": log}

    def set_routing(self, cells):
        cells = np.ascontiguousarray(cells, np.uint8)
        check(self.L.oslmi_surrogate_set_routing(self.h, cells.ctypes.data_as(ctypes.c_void_p),
                                                 cells.size))

    def set_spot_check(self, min_weight):
        check(self.L.oslmi_surrogate_set_spot_check(self.h, float(min_weight)))

    def infer(self, queries, out, material=None, fb_index=None, fb_count=None, stream=None):
        import torch
        stream = stream or torch.cuda.current_stream()
        sp = ctypes.c_void_p(stream.cuda_stream)
        m = ctypes.byref(material) if material is not None else None
        check(self.L.oslmi_surrogate_infer(self.h, m, queries.shape[1], _p(queries), _p(out),
                                           _p(fb_index) if fb_index is not None else None,
                                           _p(fb_count) if fb_count is not None else None, sp))


def route_cells(q):
    """Routing cell of every query (numpy, mirrors the kernel)."""
    q = np.asarray(q, np.float32)
    d = -q[0] * q[3] + q[2] * q[4]
    a = np.clip(((q[6] - np.float32(0.05)) * np.float32(8.0 / 0.95)).astype(np.int32), 0, 7)
    b = np.clip(((d + np.float32(1.0)) * np.float32(4.0)).astype(np.int32), 0, 7)
    return a * 8 + b


def relative_errors(pred, exact):
    """Per-query relative error of the weight (mean over rgb) and of the pdf."""
    pred = np.asarray(pred, np.float64)
    exact = np.asarray(exact, np.float64)
    den = np.maximum(np.abs(exact), 1e-30)
    ew = (np.abs(pred[:3] - exact[:3]) / den[:3]).mean(0)
    ep = np.abs(pred[3] - exact[3]) / den[3]
    zero = exact[3] == 0
    ew[zero] = np.where(np.abs(pred[:3, zero]).max(0) == 0, 0.0, 1.0)
    ep[zero] = np.where(pred[3, zero] == 0, 0.0, 1.0)
    return ew, ep


def calibrate(pred, exact, q, budget=0.01, margin=0.8, min_weight=0.0):
    """Routing table from held-out calibration queries: routing cells go to
    the exact closure, the cell with the largest mean network error first,
    until the mean relative error over all calibration queries (answers of
    the exact closure, including the spot-checked ones whose predicted
    weight is below min_weight, count as error-free) is at most
    margin * budget for both the weight and the pdf. Returns (table,
    per-cell stats, predicted mean errors (weight, pdf), predicted fallback
    fraction)."""
    ew, ep = relative_errors(pred, exact)
    spot = np.asarray(pred, np.float64)[:3].min(0) < min_weight
    ew = np.where(spot, 0.0, ew)
    ep = np.where(spot, 0.0, ep)
    cells = route_cells(q)
    n = float(len(ew))
    table = np.zeros(NCELL, np.uint8)
    stats = []
    sw = np.zeros(NCELL)
    sp = np.zeros(NCELL)
    cnt = np.zeros(NCELL)
    for c in range(NCELL):
        sel = cells == c
        cnt[c] = sel.sum()
        if cnt[c]:
            sw[c] = ew[sel].sum()
            sp[c] = ep[sel].sum()
            stats.append((c, int(cnt[c]), float(max(sw[c], sp[c]) / cnt[c])))
    mw, mp = sw.sum() / n, sp.sum() / n
    order = sorted(range(NCELL), key=lambda c: -(max(sw[c], sp[c]) / cnt[c] if cnt[c] else 0))
    for c in order:
        if max(mw, mp) <= margin * budget or not cnt[c]:
            break
        table[c] = 1
        mw -= sw[c] / n
        mp -= sp[c] / n
    routed = np.isin(cells, np.nonzero(table)[0])
    return table, stats, (float(mw), float(mp)), float((routed | spot).mean())
