"""Renderer attributes behind getattribute() (BatchedRendererServices::
get_attribute, batched_rendererservices.h:128-150): the oracle's
interpreter against the closed form of a shader that reads present,
missing and mistyped attributes, and emission + hiprtc compile of the same
shader (no GPU needed)."""
import numpy as np

import _oracle as O
from openshadinglanguage_amd import groupgen as G
from openshadinglanguage_amd import shadingsys as S

ATTRS = [("user:scale", "float", 2.5), ("user:count", "int", 7),
         ("user:tint", "color", (0.25, 0.5, 0.75)),
         ("user:xf", "matrix", np.array([[2, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [1, -1, 3, 1]],
                                        np.float32))]


def attr_shader():
    """fout = f*P.x + i + f2 + f3 + 10*(ok1 + 2 ok2 + 4 ok3 + 8 ok4 + 16 ok5
    + 32 ok6); cout = c*P + transform(m, P)."""
    o = G.Oso("shader", "attrs")
    o.sym("global", "point", "P")
    o.sym("oparam", "float", "fout", "0")
    o.sym("oparam", "color", "cout", "0 0 0")
    f = o.sym("local", "float", "f", "-1")
    f2 = o.sym("local", "float", "f2", "-3")
    f3 = o.sym("local", "float", "f3", "-5")
    i = o.sym("local", "int", "i", "-7")
    c = o.sym("local", "color", "c", "0.5 0.5 0.5")
    m = o.sym("local", "matrix", "m", "1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1")
    oks = [o.sym("local", "int", "ok%d" % k, "0") for k in range(6)]
    # locals start from explicit assignments, as oslc writes them
    for var, typ, val in ((f, "float", (-1.0,)), (f2, "float", (-3.0,)), (f3, "float", (-5.0,)),
                          (i, "int", (-7,)), (c, "color", (0.5, 0.5, 0.5))):
        o.op("assign", var, o.const(typ, *val))
    o.op("matrix", m, *[o.const("float", float(v)) for v in np.eye(4).reshape(-1)])
    for ok, (name, dst) in zip(oks, (("user:scale", f), ("user:count", i), ("user:tint", c),
                                     ("user:missing", f2), ("user:count", f3),
                                     ("user:xf", m))):
        o.op("getattribute", ok, o.const("string", name), dst)
    t = o.tmp("float")
    o.op("compref", t, "P", o.const("int", 0))
    acc = o.tmp("float")
    o.op("mul", acc, f, t)
    fi = o.tmp("float")
    o.op("assign", fi, i)
    o.op("add", acc, acc, fi)
    o.op("add", acc, acc, f2)
    o.op("add", acc, acc, f3)
    bits = o.tmp("int")
    o.op("assign", bits, o.const("int", 0))
    for k, ok in enumerate(oks):
        w = o.tmp("int")
        o.op("mul", w, ok, o.const("int", 1 << k))
        o.op("add", bits, bits, w)
    fb = o.tmp("float")
    o.op("assign", fb, bits)
    o.op("mul", fb, fb, o.const("float", 10.0))
    o.op("add", "fout", acc, fb)
    cp = o.tmp("color")
    o.op("mul", cp, c, "P")
    tp = o.tmp("point")
    o.op("transform", tp, m, "P")
    o.op("add", "cout", cp, tp)
    return o


def expected(P):
    f, i, c = 2.5, 7, np.array([0.25, 0.5, 0.75])[:, None]
    M = ATTRS[3][2].astype(np.float64)
    bits = 1 + 2 + 4 + 32  # scale, count, tint, xf found; missing and mistyped not
    fout = f * P[0] + i + (-3) + (-5) + 10 * bits
    h = np.vstack([P, np.ones((1, P.shape[1]))]).T @ M
    cout = c * P + (h[:, :3] / h[:, 3:4]).T
    return fout, cout


def test_interpreter_getattribute_closed_form():
    P = np.random.default_rng(1).uniform(-2, 2, (3, 64)).astype(np.float32)
    o = attr_shader()
    g = O.Group([(o.name, "l0", o.text())], [])
    for name, typ, v in ATTRS:
        g.set_attribute(name, typ, v)
    out, _ = g.outputs("fout", "cout").run(P, P)
    fout, cout = expected(P.astype(np.float64))
    np.testing.assert_allclose(out[0], fout, rtol=1e-6, atol=1e-5)
    np.testing.assert_allclose(out[1:4], cout, rtol=1e-6, atol=1e-5)


def test_emit_and_compile_getattribute(tmp_path):
    ss = S.ShadingSystem(jit_cache=str(tmp_path))
    for name, typ, v in ATTRS:
        ss.set_attribute(name, typ, v)
    o = attr_shader()
    g = S.build_generated_group(ss, ([(o.name, "l0", o.text())], []), "attrs")
    src = ss.emit(g)
    assert src.count("[16] == ") == 6
    ss.compile_group(g)
