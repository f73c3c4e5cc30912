"""Hand-written OSO programs that pin control flow and closure dataflow.

Each program is written as the op stream oslc produces (loops as
[init][cond][body][iter] with jump labels, codegen.cpp:1410-1445;
break/continue/return/exit as argument-less ops; user functions inlined
after a functioncall op whose jump is the op after the body), and comes
with its result in closed form, computed from what the OSL source means,
not from the op stream. The data-dependent trip counts come from the
point: n = floor(P.x), a = floor(P.y), b = floor(P.z), each in 0..8.

Used by tests/test_controlflow.py (interpreter against the closed forms)
and tests/test_gpu_controlflow.py (emitted kernels against both).
"""
import numpy as np

from openshadinglanguage_amd import groupgen as G


def _header(name):
    o = G.Oso("shader", name)
    o.sym("global", "point", "P")
    o.sym("global", "normal", "N")
    o.sym("oparam", "float", "fout", "0")
    o.sym("oparam", "color", "cout", "0 0 0")
    return o


def _ints(o):
    """n, a, b = int(floor(P.x)), int(floor(P.y)), int(floor(P.z))."""
    out = []
    for c, nm in enumerate("nab"):
        t, fl = o.tmp("float"), o.tmp("float")
        o.op("compref", t, "P", o.const("int", c))
        o.op("floor", fl, t)
        v = o.sym("local", "int", nm)
        o.op("assign", v, fl)
        out.append(v)
    return out


def _loop(o, kind, cond, init, condf, body, it):
    """A loop op (for / while / dowhile) with its four code blocks, each a
    callable that writes ops."""
    k = o.op(kind, cond)
    init()
    j0 = len(o.code)
    condf()
    j1 = len(o.code)
    body()
    j2 = len(o.code)
    it()
    j3 = len(o.code)
    o.code[k][2] = [j0, j1, j2, j3]


def _if(o, cond, then, els=lambda: None):
    k = o.op("if", cond)
    then()
    j0 = len(o.code)
    els()
    j1 = len(o.code)
    o.code[k][2] = [j0, j1]


def _nothing():
    return None


def for_continue_break():
    """for (i = 0; i < n; i += 1) { if (i == a) continue;
                                    if (i == b) break; f += i + 1; }"""
    o = _header("cf_for_continue_break")
    n, a, b = _ints(o)
    f = o.sym("local", "float", "f")
    o.op("assign", f, o.const("float", 0.0))
    i = o.sym("local", "int", "i")
    cond, ca, cb, ip, ipf = (o.tmp("int"), o.tmp("int"), o.tmp("int"), o.tmp("int"),
                             o.tmp("float"))

    def body():
        o.op("eq", ca, i, a)
        _if(o, ca, lambda: o.op("continue"))
        o.op("eq", cb, i, b)
        _if(o, cb, lambda: o.op("break"))
        o.op("add", ip, i, o.const("int", 1))
        o.op("assign", ipf, ip)
        o.op("add", f, f, ipf)

    _loop(o, "for", cond, lambda: o.op("assign", i, o.const("int", 0)),
          lambda: o.op("lt", cond, i, n), body, lambda: o.op("add", i, i, o.const("int", 1)))
    o.op("assign", "fout", f)

    def want(n, a, b):
        # the loop ends at b unless continue skips that index (a == b)
        m = np.where((b >= n) | (a == b), n, b)
        return m * (m + 1) / 2.0 - np.where(a < m, a + 1, 0)
    return o, want


def while_doubling():
    """i = 0; while (i < n) { f = f * 2 + 1; i += 1; }  ->  2^n - 1"""
    o = _header("cf_while")
    n, _, _ = _ints(o)
    f = o.sym("local", "float", "f")
    o.op("assign", f, o.const("float", 0.0))
    i = o.sym("local", "int", "i")
    o.op("assign", i, o.const("int", 0))
    cond = o.tmp("int")

    def body():
        o.op("mul", f, f, o.const("float", 2.0))
        o.op("add", f, f, o.const("float", 1.0))
        o.op("add", i, i, o.const("int", 1))

    _loop(o, "while", cond, _nothing, lambda: o.op("lt", cond, i, n), body, _nothing)
    o.op("assign", "fout", f)
    return o, lambda n, a, b: 2.0 ** n - 1


def dowhile_continue():
    """i = 0; do { i += 1; if (i == a) continue; f += i; } while (i < n);
    The body runs at least once."""
    o = _header("cf_dowhile")
    n, a, _ = _ints(o)
    f = o.sym("local", "float", "f")
    o.op("assign", f, o.const("float", 0.0))
    i = o.sym("local", "int", "i")
    o.op("assign", i, o.const("int", 0))
    cond, ca, fi = o.tmp("int"), o.tmp("int"), o.tmp("float")

    def body():
        o.op("add", i, i, o.const("int", 1))
        o.op("eq", ca, i, a)
        _if(o, ca, lambda: o.op("continue"))
        o.op("assign", fi, i)
        o.op("add", f, f, fi)

    _loop(o, "dowhile", cond, _nothing, lambda: o.op("lt", cond, i, n), body, _nothing)
    o.op("assign", "fout", f)

    def want(n, a, b):
        m = np.maximum(n, 1)
        return m * (m + 1) / 2.0 - np.where((a >= 1) & (a <= m), a, 0)
    return o, want


def nested_break():
    """for (i < n) for (j < n) { if (j == i) break; f += 1; }  ->  n(n-1)/2"""
    o = _header("cf_nested_break")
    n, _, _ = _ints(o)
    f = o.sym("local", "float", "f")
    o.op("assign", f, o.const("float", 0.0))
    i, j = o.sym("local", "int", "i"), o.sym("local", "int", "j")
    c1, c2, ce = o.tmp("int"), o.tmp("int"), o.tmp("int")

    def inner_body():
        o.op("eq", ce, j, i)
        _if(o, ce, lambda: o.op("break"))
        o.op("add", f, f, o.const("float", 1.0))

    def outer_body():
        _loop(o, "for", c2, lambda: o.op("assign", j, o.const("int", 0)),
              lambda: o.op("lt", c2, j, n), inner_body,
              lambda: o.op("add", j, j, o.const("int", 1)))

    _loop(o, "for", c1, lambda: o.op("assign", i, o.const("int", 0)),
          lambda: o.op("lt", c1, i, n), outer_body, lambda: o.op("add", i, i, o.const("int", 1)))
    o.op("assign", "fout", f)
    return o, lambda n, a, b: n * (n - 1) / 2.0


def function_return():
    """float g(int k) { if (k % 2 == 1) return 1; if (k == b) return 5;
                        return 2; }
    for (i = 0; i < n; i += 1) f += g(i);
    A return leaves the function only: the loop goes on."""
    o = _header("cf_function_return")
    n, _, b = _ints(o)
    f = o.sym("local", "float", "f")
    o.op("assign", f, o.const("float", 0.0))
    i = o.sym("local", "int", "i")
    cond, md, c1, c2 = o.tmp("int"), o.tmp("int"), o.tmp("int"), o.tmp("int")
    ret = o.tmp("float")

    def body():
        k = o.op("functioncall", o.const("string", "g"))
        o.op("mod", md, i, o.const("int", 2))
        o.op("eq", c1, md, o.const("int", 1))

        def r1():
            o.op("assign", ret, o.const("float", 1.0))
            o.op("return")
        _if(o, c1, r1)
        o.op("eq", c2, i, b)

        def r5():
            o.op("assign", ret, o.const("float", 5.0))
            o.op("return")
        _if(o, c2, r5)
        o.op("assign", ret, o.const("float", 2.0))
        o.op("return")
        o.code[k][2] = [len(o.code)]
        o.op("add", f, f, ret)

    _loop(o, "for", cond, lambda: o.op("assign", i, o.const("int", 0)),
          lambda: o.op("lt", cond, i, n), body, lambda: o.op("add", i, i, o.const("int", 1)))
    o.op("assign", "fout", f)

    def want(n, a, b):
        tot = np.zeros(np.shape(n))
        for k in range(9):
            v = np.where(k % 2 == 1, 1.0, np.where(k == b, 5.0, 2.0))
            tot += np.where(k < n, v, 0.0)
        return tot
    return o, want


def exit_early():
    """fout = 1; if (n > 4) { fout = 2; exit(); } fout = 3;"""
    o = _header("cf_exit")
    n, _, _ = _ints(o)
    c = o.tmp("int")
    o.op("assign", "fout", o.const("float", 1.0))
    o.op("gt", c, n, o.const("int", 4))

    def leave():
        o.op("assign", "fout", o.const("float", 2.0))
        o.op("exit")
    _if(o, c, leave)
    o.op("assign", "fout", o.const("float", 3.0))
    return o, lambda n, a, b: np.where(n > 4, 2.0, 3.0)


def if_else_chain():
    """if (n < a) f = 1; else if (n == a) f = 2; else { f = 3;
    if (b > 4) f = f + 10; }"""
    o = _header("cf_if_else")
    n, a, b = _ints(o)
    f = o.sym("local", "float", "f")
    c1, c2, c3 = o.tmp("int"), o.tmp("int"), o.tmp("int")
    o.op("lt", c1, n, a)

    def inner_else():
        o.op("assign", f, o.const("float", 3.0))
        o.op("gt", c3, b, o.const("int", 4))
        _if(o, c3, lambda: o.op("add", f, f, o.const("float", 10.0)))

    def els():
        o.op("eq", c2, n, a)
        _if(o, c2, lambda: o.op("assign", f, o.const("float", 2.0)), inner_else)
    _if(o, c1, lambda: o.op("assign", f, o.const("float", 1.0)), els)
    o.op("assign", "fout", f)
    return o, lambda n, a, b: np.where(n < a, 1.0, np.where(n == a, 2.0,
                                                             np.where(b > 4, 13.0, 3.0)))


PROGRAMS = {
    "for_continue_break": for_continue_break,
    "while": while_doubling,
    "dowhile_continue": dowhile_continue,
    "nested_break": nested_break,
    "function_return": function_return,
    "exit": exit_early,
    "if_else": if_else_chain,
}


# ---- closures -------------------------------------------------------------
# Every lobe below is diffuse(N): the composite of identical diffuse lobes
# with weights w_k is the diffuse BSDF scaled by sum(w_k), so eval at any
# wi above the surface returns weight sum(w_k) (BSDF * cos / pdf) and pdf
# cos(N, wi) / pi, whatever the MIS weights.

def _closure_header(name):
    o = _header(name)
    o.sym("global", "closure color", "Ci")
    return o


def closure_reuse():
    """closure color c = diffuse(N); Ci = 0.5 * c + 0.5 * c;  ->  1.0"""
    o = _closure_header("cl_reuse")
    c, h1, h2, s = (o.tmp("closure color"), o.tmp("closure color"), o.tmp("closure color"),
                    o.tmp("closure color"))
    o.op("closure", c, o.const("string", "diffuse"), "N")
    o.op("mul", h1, c, o.const("float", 0.5))
    o.op("mul", h2, c, o.const("float", 0.5))
    o.op("add", s, h1, h2)
    o.op("assign", "Ci", s)
    return o, lambda n, a, b: np.ones(np.shape(n))


def closure_loop():
    """c = diffuse(N); Ci = c; for (i < n) Ci = Ci + 0.25 * c;
    A composite holds at most 8 lobes (testrender's CompositeBSDF), so
    lobes past the eighth are dropped:  1 + 0.25 * min(n, 7)."""
    o = _closure_header("cl_loop")
    n, _, _ = _ints(o)
    c, q = o.tmp("closure color"), o.tmp("closure color")
    o.op("closure", c, o.const("string", "diffuse"), "N")
    o.op("assign", "Ci", c)
    i, cond = o.sym("local", "int", "i"), o.tmp("int")

    def body():
        o.op("mul", q, c, o.const("float", 0.25))
        o.op("add", "Ci", "Ci", q)
    _loop(o, "for", cond, lambda: o.op("assign", i, o.const("int", 0)),
          lambda: o.op("lt", cond, i, n), body, lambda: o.op("add", i, i, o.const("int", 1)))
    return o, lambda n, a, b: 1.0 + 0.25 * np.minimum(n, 7)


def closure_upstream_layers():
    """Two layers sharing Ci. Layer 0: Ci = 0.7 * diffuse(N) and
    fout = 1. Layer 1 (reads layer 0's fout as fin): if (n > 3)
    Ci += 0.3 * diffuse(N); fout = fin + 1.  ->  eval weight 0.7 or 1.0."""
    o0 = _closure_header("cl_up0")
    c, w = o0.tmp("closure color"), o0.tmp("closure color")
    o0.op("closure", c, o0.const("string", "diffuse"), "N")
    o0.op("mul", w, c, o0.const("float", 0.7))
    o0.op("assign", "Ci", w)
    o0.op("assign", "fout", o0.const("float", 1.0))
    o1 = _closure_header("cl_up1")
    o1.sym("param", "float", "fin", "0")
    n, _, _ = _ints(o1)
    cnd, c1, w1 = o1.tmp("int"), o1.tmp("closure color"), o1.tmp("closure color")
    o1.op("gt", cnd, n, o1.const("int", 3))

    def add():
        o1.op("closure", c1, o1.const("string", "diffuse"), "N")
        o1.op("mul", w1, c1, o1.const("float", 0.3))
        o1.op("add", "Ci", "Ci", w1)
    _if(o1, cnd, add)
    o1.op("add", "fout", "fin", o1.const("float", 1.0))
    return ([o0, o1], [("l0", "fout", "l1", "fin")],
            lambda n, a, b: np.where(n > 3, 1.0, 0.7))


CLOSURE_PROGRAMS = {"reuse": closure_reuse, "loop": closure_loop}


def inputs(count, seed):
    """Points with P in [0, 9)^3 (trip counts 0..8), unit normals N and
    directions wo, wi in N's upper hemisphere."""
    rng = np.random.default_rng(seed)
    P = rng.uniform(0.0, 9.0, (3, count)).astype(np.float32)
    N = rng.normal(size=(3, count))
    N /= np.linalg.norm(N, axis=0)

    def hemi():
        d = rng.normal(size=(3, count))
        d /= np.linalg.norm(d, axis=0)
        d *= np.sign(np.sum(d * N, axis=0))
        return d
    wo, wi = hemi(), hemi()
    nai = [np.floor(P[k]).astype(np.int64) for k in range(3)]
    cosi = np.sum(N.astype(np.float32).astype(np.float64) * wi.astype(np.float32), axis=0)
    return {"P": P, "N": N.astype(np.float32), "wo": wo.astype(np.float32),
            "wi": wi.astype(np.float32), "nab": nai, "cos_wi": cosi}
