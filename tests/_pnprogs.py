"""Periodic and vector noise programs (OSO, written with groupgen.Oso) for
the parity tests of the emitted kernels against the oracle's interpreter.

One shader per noise name: pnoise(name, ...) in 1-4 dimensions, each with a
float result (F1..F4) and a color result (C1..C4). Coordinates come from
the globals P and dPdu[0] (the 4th, time coordinate); periods vary per
point, from N and dPdv[0], so the lattice periods are not constants the
emitter could fold. "pnoise"/"psnoise" without a name are the unsigned and
signed periodic perlin; gabor runs periodic (pgabor / pgabor3) and, under
the plain noise op, as the vector gabor3."""
import numpy as np

from openshadinglanguage_amd import groupgen as G

# (program name, op name, noise name or None)
SPECS = [
    ("pn_perlin", "pnoise", "perlin"),
    ("pn_uperlin", "pnoise", "uperlin"),
    ("pn_cell", "pnoise", "cell"),
    ("pn_hash", "pnoise", "hash"),
    ("pn_gabor", "pnoise", "gabor"),
    ("pn_default", "pnoise", None),
    ("ps_default", "psnoise", None),
    ("n_gabor3", "noise", "gabor"),
]
OUTPUTS = ["F1", "F2", "F3", "F4", "C1", "C2", "C3", "C4"]


def program(prog, opname, noisename):
    o = G.Oso("surface", prog)
    for t, n in (("point", "P"), ("normal", "N"), ("vector", "dPdu"), ("vector", "dPdv")):
        o.sym("global", t, n)
    for k in range(1, 5):
        o.sym("oparam", "float", "F%d" % k, "0")
    for k in range(1, 5):
        o.sym("oparam", "color", "C%d" % k, "0 0 0")

    def comp(v, k):
        t = o.tmp("float")
        o.op("compref", t, v, o.const("int", k))
        return t

    px, py, pt = comp("P", 0), comp("P", 1), comp("dPdu", 0)
    nx, ny, nt = comp("N", 0), comp("N", 1), comp("dPdv", 0)
    periodic = opname != "noise"
    coords = {1: [px], 2: [px, py], 3: ["P"], 4: ["P", pt]}
    pers = {1: [nx], 2: [nx, ny], 3: ["N"], 4: ["N", nt]}
    name = [o.const("string", noisename)] if noisename else []
    for dim in (1, 2, 3, 4):
        args = coords[dim] + (pers[dim] if periodic else [])
        o.op(opname, "F%d" % dim, *(name + args))
        o.op(opname, "C%d" % dim, *(name + args))
    return o.text()


def programs():
    return [(p, program(p, op, nm)) for p, op, nm in SPECS]


def inputs(n, seed):
    """Coordinates of both signs over several lattice periods; periods mixing
    whole and fractional values and values below 1 (taken as 1)."""
    rng = np.random.default_rng(seed)
    f = np.float32
    P = rng.uniform(-12, 12, (3, n)).astype(f)
    dPdu = rng.uniform(-12, 12, (3, n)).astype(f)
    choices = np.array([0.5, 1.0, 2.0, 3.0, 3.5, 4.0, 7.25, -1.0], f)
    N = rng.choice(choices, (3, n))
    dPdv = rng.choice(choices, (3, n))
    return {"P": P, "N": N, "dPdu": dPdu, "dPdv": dPdv}
