"""MaterialX closure programs (OSO, written with groupgen.Oso) for the
parity tests of the emitted kernels and the oracle's closure walk: sums,
layers of layers, tinted and anisotropic lobes, closures reused, and the
corner cases of testrender's walks (an opaque top hides its base, a null
top leaves the base whole, a queued null base ends the walk, a zero weight
gives the null closure). Per-point parameters come from the globals P,
dPdu and dPdv."""
from openshadinglanguage_amd import groupgen as G


class _W:
    def __init__(self, name):
        o = self.o = G.Oso("surface", name)
        for t, n in (("point", "P"), ("normal", "N"), ("vector", "dPdu"), ("vector", "dPdv")):
            o.sym("global", t, n)
        o.sym("global", "closure color", "Ci")
        self.f = [self.comp("P", k) for k in range(3)]

    def comp(self, v, k):
        t = self.o.tmp("float")
        self.o.op("compref", t, v, self.o.const("int", k))
        return t

    def lin(self, x, a, b):
        """a + b * x"""
        t, u = self.o.tmp("float"), self.o.tmp("float")
        self.o.op("mul", t, self.o.const("float", b), x)
        self.o.op("add", u, self.o.const("float", a), t)
        return u

    def color(self, v):
        c = self.o.tmp("color")
        self.o.op("assign", c, v)
        return c

    def closure(self, name, *args):
        c = self.o.tmp("closure color")
        self.o.op("closure", c, self.o.const("string", name), *args)
        return c

    def zero_v(self):
        return self.o.const("vector", 0, 0, 0)

    def zero_c(self):
        return self.o.const("color", 0, 0, 0)

    def dielectric(self, rx, ry, ior, tint=None, U=None):
        return self.closure("dielectric_bsdf", "N", U or self.zero_v(),
                            tint or self.o.const("color", 1, 1, 1), self.zero_c(), rx, ry, ior,
                            self.o.const("string", "ggx"))

    def schlick(self, rx, ry, f0, f90, expo, tint=None, U=None):
        return self.closure("generalized_schlick_bsdf", "N", U or self.zero_v(),
                            tint or self.o.const("color", 1, 1, 1), self.zero_c(), rx, ry, f0, f90,
                            expo, self.o.const("string", "ggx"))

    def sheen(self, rough, albedo=None):
        return self.closure("sheen_bsdf", "N", albedo or self.o.const("color", 1, 1, 1), rough)

    def diffuse(self):
        return self.closure("diffuse", "N")

    def ggx(self, alpha):
        return self.closure("microfacet", self.o.const("string", "ggx"), "N", self.zero_v(), alpha,
                            alpha, self.o.const("float", 1.5), self.o.const("int", 0))

    def ggx_aniso(self, U, xalpha, yalpha, dist="ggx", refract=0, eta=1.5):
        return self.closure("microfacet", self.o.const("string", dist), "N", U, xalpha, yalpha,
                            self.o.const("float", eta), self.o.const("int", refract))

    def layer(self, top, base):
        return self.closure("layer", top, base)

    def add(self, a, b):
        c = self.o.tmp("closure color")
        self.o.op("add", c, a, b)
        return c

    def mul(self, w, c):
        r = self.o.tmp("closure color")
        self.o.op("mul", r, w, c)
        return r

    def fconst(self, v):
        return self.o.const("float", v)

    def cconst(self, *v):
        return self.o.const("color", *v)

    def done(self, ci):
        self.o.op("assign", "Ci", ci)
        return self.o


def programs():
    """[(name, Oso)]: every program's Ci built from MaterialX closures."""
    out = []

    # a plain sum: anisotropic dielectric (tangent dPdu), tinted Schlick,
    # coloured sheen
    w = _W("mx_sum")
    x, y, z = w.f
    d = w.dielectric(w.lin(x, 0.05, 0.6), w.lin(y, 0.1, 0.8), w.lin(z, 1.3, 0.5), U="dPdu")
    s = w.schlick(w.lin(y, 0.05, 0.9), w.lin(y, 0.05, 0.9), w.color("dPdu"), w.color("dPdv"),
                  w.fconst(5.0), tint=w.cconst(0.9, 0.7, 0.5))
    sh = w.sheen(w.lin(z, 0.1, 0.9), albedo=w.color("dPdv"))
    out.append(("mx_sum", w.done(w.add(w.add(w.mul(w.fconst(0.6), d), w.mul(w.color("P"), s)), sh))))

    # a dielectric coat over a diffuse base, the whole scaled
    w = _W("mx_coat_diffuse")
    x, y, z = w.f
    out.append(("mx_coat_diffuse", w.done(w.mul(w.cconst(0.8, 0.9, 1.0), w.layer(
        w.dielectric(w.lin(y, 0.05, 0.9), w.lin(y, 0.05, 0.9), w.lin(x, 1.3, 0.5)), w.diffuse())))))

    # layer of layers: the top is itself a layer (its opacity is its base's)
    w = _W("mx_nested")
    x, y, z = w.f
    top = w.layer(w.sheen(w.lin(z, 0.1, 0.9)), w.dielectric(w.lin(y, 0.05, 0.9), w.lin(y, 0.05, 0.9),
                                                            w.lin(x, 1.3, 0.5)))
    base = w.schlick(w.lin(x, 0.05, 0.9), w.lin(x, 0.05, 0.9), w.color("dPdu"), w.color("dPdv"),
                     w.lin(z, 1.0, 6.0))
    out.append(("mx_nested", w.done(w.layer(top, base))))

    # a sum as the top: its opacity is that of the last closure walked
    w = _W("mx_sum_top")
    x, y, z = w.f
    top = w.add(w.sheen(w.lin(z, 0.1, 0.9), albedo=w.cconst(0.5, 0.6, 0.7)),
                w.dielectric(w.lin(y, 0.05, 0.9), w.lin(y, 0.05, 0.9), w.lin(x, 1.3, 0.5)))
    out.append(("mx_sum_top", w.done(w.add(w.layer(top, w.ggx(w.lin(y, 0.1, 0.5))),
                                           w.mul(w.fconst(0.25), w.diffuse())))))

    # an opaque top (diffuse) hides the base
    w = _W("mx_opaque_top")
    x, y, z = w.f
    out.append(("mx_opaque_top", w.done(w.layer(w.diffuse(), w.sheen(w.lin(z, 0.1, 0.9))))))

    # a null top (zero weight) leaves the base whole
    w = _W("mx_null_top")
    x, y, z = w.f
    out.append(("mx_null_top", w.done(w.layer(w.mul(w.fconst(0.0), w.sheen(w.lin(z, 0.1, 0.9))),
                                              w.schlick(w.lin(y, 0.05, 0.9), w.lin(y, 0.05, 0.9),
                                                        w.color("dPdu"), w.color("dPdv"),
                                                        w.fconst(5.0))))))

    # a queued null base ends the walk: the diffuse added after it is never
    # visited
    w = _W("mx_null_base")
    x, y, z = w.f
    lay = w.layer(w.sheen(w.lin(z, 0.1, 0.9)), w.mul(w.fconst(0.0), w.diffuse()))
    out.append(("mx_null_base", w.done(w.add(lay, w.diffuse()))))

    # one closure value used three times (capacity follows the dataflow)
    w = _W("mx_reuse")
    x, y, z = w.f
    c = w.dielectric(w.lin(y, 0.05, 0.9), w.lin(x, 0.05, 0.9), w.lin(z, 1.3, 0.5))
    out.append(("mx_reuse", w.done(w.add(w.layer(c, c), w.mul(w.fconst(0.5), c)))))

    # a dielectric coat over an anisotropic GGX base (tangent dPdu, the two
    # roughnesses from P), plus an isotropic GGX lobe in the same list
    w = _W("mx_coat_aniso")
    x, y, z = w.f
    base = w.ggx_aniso("dPdu", w.lin(x, 0.05, 0.5), w.lin(y, 0.05, 0.5))
    coat = w.dielectric(w.lin(z, 0.05, 0.5), w.lin(z, 0.05, 0.5), w.fconst(1.5))
    out.append(("mx_coat_aniso", w.done(w.add(w.layer(coat, base),
                                              w.mul(w.fconst(0.3), w.ggx(w.lin(z, 0.1, 0.4)))))))

    # glass: a sheen over a rough anisotropic GGX that reflects and refracts
    # (refract 2), plus a refract-only isotropic lobe from the inside
    # (eta < 1)
    w = _W("mx_glass")
    x, y, z = w.f
    glass = w.ggx_aniso("dPdv", w.lin(x, 0.05, 0.4), w.lin(y, 0.05, 0.4), refract=2, eta=1.45)
    inner = w.ggx_aniso(w.zero_v(), w.lin(z, 0.1, 0.3), w.lin(z, 0.1, 0.3), refract=1,
                        eta=1.0 / 1.33)
    out.append(("mx_glass", w.done(w.add(w.layer(w.sheen(w.lin(z, 0.1, 0.5)), glass),
                                         w.mul(w.fconst(0.5), inner)))))

    # Beckmann lobes, sampled as well: an anisotropic one under a sheen and
    # an isotropic one beside it
    w = _W("mx_beckmann")
    x, y, z = w.f
    bk = w.ggx_aniso("dPdu", w.lin(x, 0.05, 0.5), w.lin(y, 0.05, 0.5), dist="beckmann")
    iso = w.ggx_aniso(w.zero_v(), w.lin(z, 0.1, 0.4), w.lin(z, 0.1, 0.4), dist="beckmann")
    out.append(("mx_beckmann", w.done(w.add(w.layer(w.sheen(w.lin(z, 0.1, 0.5)), bk),
                                            w.mul(w.fconst(0.4), iso)))))

    # Phong under a sheen, plus a translucent lobe (diffuse about -N)
    w = _W("mx_phong")
    x, y, z = w.f
    ph = w.closure("phong", "N", w.lin(x, 2.0, 60.0))
    out.append(("mx_phong", w.done(w.add(w.layer(w.sheen(w.lin(z, 0.1, 0.5)), ph),
                                         w.mul(w.fconst(0.3), w.closure("translucent", "N"))))))

    # delta lobes beside a diffuse one: a Fresnel mirror under a sheen, a
    # perfect mirror, a refraction and a transparent lobe
    w = _W("mx_delta")
    x, y, z = w.f
    fres = w.closure("reflection", "N", w.lin(x, 1.2, 0.6))
    out.append(("mx_delta", w.done(w.add(
        w.add(w.layer(w.sheen(w.lin(z, 0.1, 0.5)), fres), w.mul(w.fconst(0.2), w.closure("reflection", "N"))),
        w.add(w.mul(w.fconst(0.3), w.closure("refraction", "N", w.lin(y, 1.1, 0.5))),
              w.add(w.mul(w.fconst(0.1), w.closure("transparent")), w.mul(w.fconst(0.4), w.diffuse())))))))

    # emission beside a BSDF: exported with its weight, ignored by eval and
    # sampling
    w = _W("mx_emission")
    x, y, z = w.f
    out.append(("mx_emission", w.done(w.add(w.mul(w.color("dPdu"), w.closure("emission")),
                                            w.layer(w.sheen(w.lin(z, 0.1, 0.5)), w.diffuse())))))
    return out


def equivalents():
    """{program: Oso of a simpler closure the walk rules make it equal to}"""
    eq = {}
    w = _W("eq_opaque_top")
    w.f
    eq["mx_opaque_top"] = w.done(w.diffuse())
    w = _W("eq_null_top")
    x, y, z = w.f
    eq["mx_null_top"] = w.done(w.schlick(w.lin(y, 0.05, 0.9), w.lin(y, 0.05, 0.9), w.color("dPdu"),
                                         w.color("dPdv"), w.fconst(5.0)))
    w = _W("eq_null_base")
    x, y, z = w.f
    eq["mx_null_base"] = w.done(w.sheen(w.lin(z, 0.1, 0.9)))
    return eq


def inputs(n, seed):
    """Globals and directions: P in [0,1)^3 (the parameters), dPdu / dPdv
    in [0,1)^3 (F0, F90, tints; dPdu also a tangent), N on the sphere, wo
    cosine-distributed about N, wi on the sphere, rnd in [0,1)^3."""
    import numpy as np
    from openshadinglanguage_amd import workloads as W
    w = W.layered_closure_inputs(n=n, seed=seed)
    rng = np.random.default_rng(seed + 1)
    f = lambda a: np.ascontiguousarray(a, np.float32)  # noqa: E731
    return {"P": f(rng.uniform(0, 1, (3, n))), "dPdu": f(rng.uniform(0, 1, (3, n))),
            "dPdv": f(rng.uniform(0, 1, (3, n))), "N": w["N"], "wo": w["V"], "wi": w["L"],
            "rnd": f(np.concatenate([w["rnd"][:2], np.zeros((1, n), np.float32)]))}
