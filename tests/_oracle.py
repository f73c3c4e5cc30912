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


def pnoise(name, outdim, inp, period, derivs=False):
    """Periodic noise. inp as for noise(); period: float32 [indim, n]
    (one period per input dimension and point)."""
    L = lib()
    L.oref_pnoise.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    L.oref_pnoise.restype = ctypes.c_int
    inp = np.ascontiguousarray(inp, dtype=np.float32)
    planes, n = inp.shape
    indim = planes // 3 if derivs else planes
    per = np.ascontiguousarray(np.broadcast_to(np.asarray(period, np.float32), (indim, n)))
    out = np.zeros(((3 if derivs else 1) * outdim, n), dtype=np.float32)
    rc = L.oref_pnoise(name.encode(), outdim, indim, int(derivs), _p(inp), _p(per), _p(out), n)
    if rc != 0:
        raise ValueError("oracle rejected periodic noise request %r" % name)
    return out


def perlin_cgdefault(outdim, inp, derivs=False):
    """Signed perlin in the CGDefault (SIMD) interpolation order."""
    L = lib()
    L.oref_perlin_cgdefault.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    inp = np.ascontiguousarray(inp, dtype=np.float32)
    indim = inp.shape[0] // (3 if derivs else 1)
    out = np.zeros(((3 if derivs else 1) * outdim, inp.shape[1]), np.float32)
    assert L.oref_perlin_cgdefault(outdim, indim, 1 if derivs else 0, _p(inp), _p(out),
                                   inp.shape[1]) == 0
    return out


def hash_op(inp):
    inp = np.ascontiguousarray(inp, dtype=np.float32)
    out = np.zeros(inp.shape[1], dtype=np.int32)
    assert lib().oref_hash_op(inp.shape[0], _p(inp), _p(out), inp.shape[1]) == 0
    return out


def _setup_more():
    L = lib()
    P = ctypes.c_void_p
    L.oref_gabor.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, ctypes.c_float,
                             ctypes.c_float, P, P, ctypes.c_int64]
    L.oref_gabor.restype = ctypes.c_int
    L.oref_blackbody.argtypes = [P, P, ctypes.c_int64]
    L.oref_colorsys.argtypes = [ctypes.c_char_p, P]
    L.oref_shade_c1.argtypes = [P, P, ctypes.c_int64, ctypes.c_int]
    L.oref_shade_c2.argtypes = [P, P, P, P, P, P, P, P, ctypes.c_int64, ctypes.c_int]
    return L


def gabor(inp, anisotropic=0, do_filter=1, direction=(1.0, 0.0, 0.0), bandwidth=1.0,
          impulses=16.0):
    """inp: [3*dim, n] Dual2 planes. Returns [3, n] (val, dx, dy)."""
    L = _setup_more()
    inp = np.ascontiguousarray(inp, dtype=np.float32)
    dim = inp.shape[0] // 3
    out = np.zeros((3, inp.shape[1]), np.float32)
    d = np.asarray(direction, np.float32)
    assert L.oref_gabor(dim, anisotropic, do_filter, _p(d), bandwidth, impulses, _p(inp), _p(out),
                        inp.shape[1]) == 0
    return out


def gabor_ex(inp, outdim=1, period=None, anisotropic=0, do_filter=1, direction=(1.0, 0.0, 0.0),
             bandwidth=1.0, impulses=16.0):
    """Gabor noise of 1 or 3 components (gabor / gabor3), periodic when
    period ([dim, n] or broadcastable) is given. inp: [3*dim, n] Dual2
    planes. Returns [3*outdim, n] (values, dx, dy)."""
    L = _setup_more()
    P = ctypes.c_void_p
    L.oref_gabor_ex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P,
                                ctypes.c_float, ctypes.c_float, P, P, P, ctypes.c_int64]
    L.oref_gabor_ex.restype = ctypes.c_int
    inp = np.ascontiguousarray(inp, dtype=np.float32)
    dim, n = inp.shape[0] // 3, inp.shape[1]
    per = None
    if period is not None:
        per = np.ascontiguousarray(np.broadcast_to(np.asarray(period, np.float32), (dim, n)))
    out = np.zeros((3 * outdim, n), np.float32)
    d = np.asarray(direction, np.float32)
    assert L.oref_gabor_ex(dim, outdim, anisotropic, do_filter, _p(d), bandwidth, impulses,
                           _p(inp), None if per is None else _p(per), _p(out), n) == 0
    return out


def blackbody(T):
    L = _setup_more()
    T = np.ascontiguousarray(T, dtype=np.float32)
    out = np.zeros((3, T.size), np.float32)
    L.oref_blackbody(_p(T), _p(out), T.size)
    return out


def transformc(frm, to, c):
    """ColorSystem::transformc(from, to, c) over 3 planes (to None: the
    color(space, ...) constructor's to_rgb)."""
    L = lib()
    L.oref_transformc.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_int64]
    c = np.ascontiguousarray(c, np.float32).reshape(3, -1)
    out = np.zeros_like(c)
    rc = L.oref_transformc(frm.encode(), None if to is None else to.encode(), _p(c), _p(out),
                           c.shape[1])
    if rc != 0:
        raise ValueError("color space needs OpenColorIO: %s -> %s" % (frm, to))
    return out


def colorsys(name="Rec709"):
    L = _setup_more()
    out = np.zeros(973, np.float32)
    L.oref_colorsys(name.encode(), _p(out))
    return {"bb_table": out[:951].reshape(317, 3), "xyz2rgb": out[951:960].reshape(3, 3),
            "rgb2xyz": out[960:969].reshape(3, 3), "lum": out[969:972],
            "bb_entries": int(out[972])}


def shade_c1(P, nthreads=1):
    L = _setup_more()
    P = np.ascontiguousarray(P, dtype=np.float32)
    out = np.zeros_like(P)
    L.oref_shade_c1(_p(P), _p(out), P.shape[1], nthreads)
    return out


def shade_c2(P, N, wo, wi, dPdx, dPdy, nthreads=1, debug=False):
    L = _setup_more()
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (P, N, wo, wi)]
    n = arrs[0].shape[1]
    dx = np.ascontiguousarray(dPdx, np.float32)
    dy = np.ascontiguousarray(dPdy, np.float32)
    out = np.zeros((4, n), np.float32)
    dbg = np.zeros((20, n), np.float32) if debug else None
    L.oref_shade_c2(*[_p(a) for a in arrs], _p(dx), _p(dy), _p(out),
                    _p(dbg) if debug else None, n, nthreads)
    return (out, dbg) if debug else out


def omath(fn, x):
    """The oracle's fixed transcendentals: 'exp', 'sin', 'cos', 'cbrt', 'log',
    'erf', 'ierf'."""
    L = lib()
    L.oref_math.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros_like(x)
    L.oref_math(["exp", "sin", "cos", "cbrt", "log", "erf", "ierf"].index(fn), _p(x), _p(out),
                x.size)
    return out


# ---- layered MaterialX closure (oracle/mtx.cpp) ----------------------------
MTX_FN = {"ggx_D": (0, 5, 1), "ggx_G1": (1, 5, 1), "ggx_G2_G1": (2, 8, 1),
          "ggx_D_refl_D": (3, 5, 1), "vndf": (4, 7, 3), "bvndf": (5, 7, 3),
          "unit_disc": (6, 2, 2), "uniform_hemisphere": (7, 2, 3), "fresnel": (8, 2, 1),
          "emiss": (9, 4, 1), "ck_albedo": (10, 2, 1), "ck_D": (11, 4, 1), "pow": (12, 2, 1),
          "log2": (13, 1, 1), "exp2": (14, 1, 1)}


def _mtx_setup():
    L = lib()
    P, I64 = ctypes.c_void_p, ctypes.c_int64
    L.oref_mtx_layered.argtypes = [P, P, P, P, P, P, P, P, I64, ctypes.c_int]
    L.oref_mtx_eval_local.argtypes = [P, P, P, I64, ctypes.c_int]
    L.oref_mtx_weights.argtypes = [P, P, P, P, P, P, I64]
    L.oref_mtx_luts.argtypes = [P]
    L.oref_mtx_bake.argtypes = [P, ctypes.c_int]
    L.oref_mtx_fn.argtypes = [ctypes.c_int, P, P, I64]
    L.oref_mtx_fn.restype = ctypes.c_int
    return L


def mtx_fn(name, rows):
    """Building block `name` on AoS rows [n, width_in] -> [n, width_out]."""
    fn, win, wout = MTX_FN[name]
    rows = np.ascontiguousarray(np.asarray(rows, np.float32).reshape(-1, win))
    out = np.zeros((rows.shape[0], wout), np.float32)
    assert _mtx_setup().oref_mtx_fn(fn, _p(rows), _p(out), rows.shape[0]) == 0
    return out


def mtx_tables():
    out = np.zeros(8448, np.float32)
    _mtx_setup().oref_mtx_luts(_p(out))
    return out


def mtx_bake(nthreads=8):
    out = np.zeros(8448, np.float32)
    _mtx_setup().oref_mtx_bake(_p(out), nthreads)
    return out


def mtx_layered(mat, N, V, L, params, rnd, nthreads=1, eval_=True, samples=True):
    """Returns (eval [4, n] or None, samples [28, n] or None)."""
    Lb = _mtx_setup()
    m = np.ascontiguousarray(mat, np.float32)
    a = [np.ascontiguousarray(x, np.float32) for x in (N, V, L, params, rnd)]
    n = a[0].shape[1]
    ev = np.zeros((4, n), np.float32) if eval_ else None
    sp = np.zeros((28, n), np.float32) if samples else None
    Lb.oref_mtx_layered(_p(m), *[_p(x) for x in a], _p(ev) if eval_ else None,
                        _p(sp) if samples else None, n, nthreads)
    return ev, sp


def mtx_eval_local(mat, q, nthreads=1):
    Lb = _mtx_setup()
    m = np.ascontiguousarray(mat, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    out = np.zeros((4, q.shape[1]), np.float32)
    Lb.oref_mtx_eval_local(_p(m), _p(q), _p(out), q.shape[1], nthreads)
    return out


def mtx_weights(mat, N, V, params):
    Lb = _mtx_setup()
    m = np.ascontiguousarray(mat, np.float32)
    a = [np.ascontiguousarray(x, np.float32) for x in (N, V, params)]
    n = a[0].shape[1]
    w9 = np.zeros((9, n), np.float32)
    p3 = np.zeros((3, n), np.float32)
    Lb.oref_mtx_weights(_p(m), *[_p(x) for x in a], _p(w9), _p(p3), n)
    return w9, p3


# ---- OSO group interpreter (oracle/osoint.cpp) ---------------------------------
class Group:
    """A shader group run by the oracle's OSO interpreter."""

    def __init__(self, layers, conns):
        L = lib()
        P, C = ctypes.c_void_p, ctypes.c_char_p
        L.oref_group_new.restype = P
        L.oref_group_free.argtypes = [P]
        L.oref_group_layer.argtypes = [P, C, C]
        L.oref_group_connect.argtypes = [P, C, C, C, C]
        L.oref_group_run.argtypes = [P, P, P, P, P, ctypes.c_int64, P, P, ctypes.c_int]
        self.L = L
        self.h = L.oref_group_new()
        for shadername, layername, text in layers:
            assert L.oref_group_layer(self.h, text.encode(), layername.encode()) == 0
        for s, sp, d, dp in conns:
            assert L.oref_group_connect(self.h, s.encode(), sp.encode(), d.encode(),
                                        dp.encode()) == 0

    def __del__(self):
        try:
            self.L.oref_group_free(self.h)
        except Exception:
            pass

    def name_transform(self, name, matrix, inverse=None):
        """A named coordinate system (to-common matrix, optional inverse)."""
        self.L.oref_group_name_transform.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                                     ctypes.c_void_p, ctypes.c_void_p]
        m = np.ascontiguousarray(np.asarray(matrix, np.float32).reshape(16))
        inv = None if inverse is None else np.ascontiguousarray(
            np.asarray(inverse, np.float32).reshape(16))
        self.L.oref_group_name_transform(self.h, name.encode(), _p(m),
                                         None if inv is None else _p(inv))
        return self

    def set_attribute(self, name, typ, value):
        """A renderer attribute for getattribute()."""
        self.L.oref_group_set_attribute.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                                    ctypes.c_int, ctypes.c_void_p]
        base = {"int": 0, "float": 1, "matrix": 3}.get(typ, 2)
        v = np.ascontiguousarray(np.asarray(value, np.int32 if typ == "int" else np.float32)
                                 .reshape(-1))
        self.L.oref_group_set_attribute(self.h, name.encode(), base, _p(v))
        return self

    def outputs(self, fname, cname):
        """Which last-layer float / triple outputs run() reports."""
        self.L.oref_group_outputs.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]
        self.L.oref_group_outputs(self.h, fname.encode(), cname.encode())
        return self

    def run(self, P, N, wo=None, wi=None, nthreads=1):
        a = [np.ascontiguousarray(x, np.float32) for x in (P, N)]
        n = a[0].shape[1]
        out = np.zeros((4, n), np.float32)
        ev = np.zeros((4, n), np.float32) if wo is not None else None
        w = [np.ascontiguousarray(x, np.float32) for x in (wo, wi)] if wo is not None else None
        self.L.oref_group_run(self.h, _p(a[0]), _p(a[1]), _p(w[0]) if w else None,
                              _p(w[1]) if w else None, n, _p(out), _p(ev) if w else None,
                              nthreads)
        return out, ev


def composite_eval(lobes, wo, wi):
    """CompositeBSDF prepare + eval of one lobe list at n direction pairs.
    lobes: rows (kind, w_r, w_g, w_b, Nx, Ny, Nz, alpha, eta), optionally
    followed by (Ux, Uy, Uz, yalpha) for an anisotropic microfacet lobe;
    kind 0 diffuse, 1 GGX reflection, 2 Beckmann. Returns 4 planes (weight
    rgb, pdf)."""
    L = lib()
    L.oref_composite_eval.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_int]
    lo = np.ascontiguousarray(np.atleast_2d(lobes), np.float32)
    a = [np.ascontiguousarray(x, np.float32) for x in (wo, wi)]
    n = a[0].shape[1]
    out = np.zeros((4, n), np.float32)
    L.oref_composite_eval(lo.shape[0], _p(lo), _p(a[0]), _p(a[1]), _p(out), n, lo.shape[1])
    return out


# ---- FP64 evaluation build (oracle/f64/f64_main.cpp) -----------------------
LIB64 = os.path.join(ROOT, "oracle", "build", "liboslref_f64.so")
_lib64 = None


def lib64():
    """The oracle compiled in binary64 with exact transcendentals: the
    mathematics of the shading path, against which binary32 results get an
    error. Inputs are the binary32 inputs widened exactly."""
    global _lib64
    if _lib64 is None:
        if not os.path.exists(LIB64):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "f64"])
        L = ctypes.CDLL(LIB64)
        P, I64, I = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        L.oref64_shade_c1.argtypes = [P, P, I64, I]
        L.oref64_shade_c2.argtypes = [P] * 7 + [I64, I]
        L.oref64_mtx_layered.argtypes = [P] * 8 + [I64, I]
        L.oref64_math.argtypes = [I, P, P, I64]
        _lib64 = L
    return _lib64


def _d(a):
    return np.ascontiguousarray(np.asarray(a, np.float32), np.float64)


def shade_c1_f64(P, nthreads=1):
    P = _d(P)
    out = np.zeros_like(P)
    lib64().oref64_shade_c1(_p(P), _p(out), P.shape[1], nthreads)
    return out


def shade_c2_f64(P, N, wo, wi, dPdx, dPdy, nthreads=1):
    a = [_d(x) for x in (P, N, wo, wi, dPdx, dPdy)]
    n = a[0].shape[1]
    out = np.zeros((4, n), np.float64)
    lib64().oref64_shade_c2(*[_p(x) for x in a], _p(out), n, nthreads)
    return out


def mtx_layered_f64(mat, N, V, L, params, rnd, nthreads=1):
    """Returns (eval [4, n], samples [28, n]) in binary64."""
    a = [_d(x) for x in (mat, N, V, L, params, rnd)]
    n = a[1].shape[1]
    ev = np.zeros((4, n), np.float64)
    sp = np.zeros((28, n), np.float64)
    lib64().oref64_mtx_layered(*[_p(x) for x in a], _p(ev), _p(sp), n, nthreads)
    return ev, sp


def omath_f64(fn, x):
    """The FP64 build's exp/sin/cos/cbrt (the exact binary64 functions)."""
    x = _d(x)
    out = np.zeros_like(x)
    lib64().oref64_math(["exp", "sin", "cos", "cbrt"].index(fn), _p(x), _p(out), x.size)
    return out


def f32_error(got, ref64):
    """Error of binary32 results against a binary64 evaluation, per element:
    (relative error, error in units of the binary32 spacing at the exact
    value). Non-finite pairs count as exact only when both sides agree."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref64, np.float64)
    d = np.abs(got - ref)
    both = ~np.isfinite(got) | ~np.isfinite(ref)
    d[both] = np.where(got[both] == ref[both], 0.0, np.inf)
    mag = np.abs(ref)
    rel = d / np.maximum(mag, 1e-30)
    e = np.floor(np.log2(np.maximum(mag, np.float64(np.finfo(np.float32).tiny))))
    ulp = d / np.exp2(e - 23)
    return rel, ulp


def error_summary(got, ref64):
    """Percentiles of the binary32-vs-binary64 error over all elements."""
    rel, ulp = f32_error(got, ref64)
    rel, ulp = rel.ravel(), ulp.ravel()
    q = [50, 99, 99.9]
    s = {"n": int(rel.size)}
    for name, v in (("rel", rel), ("ulp", ulp)):
        fin = v[np.isfinite(v)]
        for p, x in zip(q, np.percentile(fin, q)):
            s["%s_p%s" % (name, str(p).replace(".", "_"))] = float(x)
        s[name + "_max"] = float(fin.max())
        s[name + "_mean"] = float(fin.mean())
    s["frac_ulp_le4"] = float(np.mean(ulp <= 4.0))
    s["frac_rel_le1e5"] = float(np.mean((ulp <= 4.0) | (rel <= 1e-5)))
    return s


def sample_flips(samp, samp64, tol=1e-3):
    """Per (sample, point) of a [28, n] sample block (4 samples x wi xyz,
    weight rgb, pdf): True where binary32 rounding changed a discrete
    choice (lobe, reflection or refraction, rejection), seen as a component
    that moved by more than tol relative to max(|exact|, 1)."""
    a = np.asarray(samp, np.float64).reshape(4, 7, -1)
    b = np.asarray(samp64, np.float64).reshape(4, 7, -1)
    d = np.abs(a - b) / np.maximum(np.abs(b), 1.0)
    d[~np.isfinite(d)] = np.inf
    return np.any(d > tol, axis=1)


def sample_error_summary(samp, samp64):
    """error_summary over the samples whose discrete choices agree, plus
    the fraction of samples whose choices differ."""
    flip = sample_flips(samp, samp64)
    keep = np.broadcast_to(~flip[:, None, :], (4, 7, flip.shape[1])).reshape(28, -1)
    s = error_summary(np.asarray(samp)[keep], np.asarray(samp64)[keep])
    s["frac_flipped"] = float(flip.mean())
    return s


def composite_sample(lobes, wo, rnd):
    """CompositeBSDF prepare + sample of one lobe list (rows as
    composite_eval) at n (wo, rnd) pairs. Returns 7 planes (wi xyz, weight
    rgb, pdf)."""
    L = lib()
    L.oref_composite_sample.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.c_int]
    lo = np.ascontiguousarray(np.atleast_2d(lobes), np.float32)
    a = [np.ascontiguousarray(x, np.float32) for x in (wo, rnd)]
    n = a[0].shape[1]
    out = np.zeros((7, n), np.float32)
    L.oref_composite_sample(lo.shape[0], _p(lo), _p(a[0]), _p(a[1]), _p(out), n, lo.shape[1])
    return out


def group_run_ex(group, P, N, wo, dPdu=None, dPdv=None, wi=None, rnd=None, nthreads=1):
    """The interpreter group's outputs, with the globals dPdu / dPdv, and
    Ci evaluated at (wo, wi) and / or sampled at (wo, rnd): (out4, eval4 or
    None, samp7 or None)."""
    L = group.L
    L.oref_group_run_ex.argtypes = [ctypes.c_void_p] * 8 + [ctypes.c_int64] + [ctypes.c_void_p] * 3 + [
        ctypes.c_int]
    f = lambda x: None if x is None else np.ascontiguousarray(x, np.float32)  # noqa: E731
    P, N, wo, dPdu, dPdv, wi, rnd = map(f, (P, N, wo, dPdu, dPdv, wi, rnd))
    n = P.shape[1]
    out = np.zeros((4, n), np.float32)
    ev = np.zeros((4, n), np.float32) if wi is not None else None
    samp = np.zeros((7, n), np.float32) if rnd is not None else None
    q = lambda x: None if x is None else _p(x)  # noqa: E731
    L.oref_group_run_ex(group.h, _p(P), _p(N), q(dPdu), q(dPdv), _p(wo), q(wi), q(rnd), n, _p(out),
                        q(ev), q(samp), nthreads)
    return out, ev, samp


def group_closures(group, P, N, wo, dPdu=None, dPdv=None, max_lobes=8):
    """Ci's lobes in the closure-export layout: [1 + 25 max_lobes, n]."""
    L = group.L
    L.oref_group_closures.argtypes = [ctypes.c_void_p] * 6 + [ctypes.c_int64, ctypes.c_int,
                                                              ctypes.c_void_p]
    f = lambda x: None if x is None else np.ascontiguousarray(x, np.float32)  # noqa: E731
    P, N, wo, dPdu, dPdv = map(f, (P, N, wo, dPdu, dPdv))
    n = P.shape[1]
    out = np.zeros((1 + 25 * max_lobes, n), np.float32)
    q = lambda x: None if x is None else _p(x)  # noqa: E731
    L.oref_group_closures(group.h, _p(P), _p(N), q(dPdu), q(dPdv), _p(wo), n, max_lobes, _p(out))
    return out


def group_run_sample(group, P, N, wo, rnd, nthreads=1):
    """The interpreter group's outputs and Ci sampled at (wo, rnd)."""
    L = group.L
    L.oref_group_run_sample.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int64, ctypes.c_void_p,
                                                               ctypes.c_void_p, ctypes.c_int]
    a = [np.ascontiguousarray(x, np.float32) for x in (P, N, wo, rnd)]
    n = a[0].shape[1]
    out = np.zeros((4, n), np.float32)
    samp = np.zeros((7, n), np.float32)
    L.oref_group_run_sample(group.h, *[_p(x) for x in a], n, _p(out), _p(samp), nthreads)
    return out, samp
