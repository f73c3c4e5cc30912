"""GPU parity of color-space transforms in emitted kernels: transformc over
rgb / hsv / hsl / YIQ / XYZ / xyY / sRGB pairs and the color(space, ...)
constructor, against the oracle's interpreter (stated tolerance 4 ulp or
1e-5 relative; both evaluate the same binary32 operations, sRGB's pow
included), and their Dual2 derivatives against central differences of the
value path and the reference's golden derivative lines
(testsuite/transformc/ref/out.txt)."""
import json
import os

import numpy as np
import pytest

import _oracle as O
from openshadinglanguage_amd import groupgen as G
from openshadinglanguage_amd import shadingsys as S

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PAIRS = ([("rgb", s) for s in ("hsv", "hsl", "YIQ", "XYZ", "xyY", "sRGB")]
         + [(s, "rgb") for s in ("hsv", "hsl", "YIQ", "XYZ", "xyY", "sRGB")]
         + [("hsv", "hsl"), ("hsl", "hsv"), ("xyY", "YIQ"), ("sRGB", "XYZ"), ("XYZ", "xyY"),
            ("YIQ", "sRGB"), ("hsv", "hsv"), ("linear", "sRGB")])
CTOR = ["hsv", "hsl", "YIQ", "XYZ", "xyY"]


def ulp_close(got, ref, ulps=4, rel=1e-5):
    got = np.asarray(got, np.float32)
    ref = np.asarray(ref, np.float32)
    gi = got.view(np.int32).astype(np.int64)
    ri = ref.view(np.int32).astype(np.int64)
    gi = np.where(gi < 0, -(gi & 0x7fffffff), gi)
    ri = np.where(ri < 0, -(ri & 0x7fffffff), ri)
    d = np.abs(got.astype(np.float64) - ref)
    return (np.abs(gi - ri) <= ulps) | (d <= rel * np.abs(ref.astype(np.float64))) | (got == ref)


def pair_shader(name, pairs, ctor=False):
    """cout_k = transformc(from_k, to_k, color(P)), or color(space_k, P.x,
    P.y, P.z) with ctor, for each k."""
    o = G.Oso("shader", name)
    o.sym("global", "point", "P")
    c = o.tmp("color")
    o.op("assign", c, "P")
    comps = []
    if ctor:
        for k in range(3):
            t = o.tmp("float")
            o.op("compref", t, "P", o.const("int", k))
            comps.append(t)
    outs = []
    for k, pr in enumerate(pairs):
        out = o.sym("oparam", "color", "c%d" % k, "0 0 0")
        if ctor:
            o.op("color", out, o.const("string", pr), *comps)
        else:
            o.op("transformc", out, o.const("string", pr[0]), o.const("string", pr[1]), c)
        outs.append(out)
    return o, outs


@pytest.fixture(scope="module")
def dev():
    from openshadinglanguage_amd import check, lib
    check(lib().oslmi_init())
    return torch.device("cuda:0")


def run(dev, o, outs, P, dPdx=(0.0, 0.0, 0.0), dPdy=(0.0, 0.0, 0.0)):
    ss = S.ShadingSystem()
    g = S.build_generated_group(ss, ([(o.name, "l0", o.text())], []), o.name, outputs=outs)
    ss.jit_group(g)
    n = P.shape[1]
    print("color group %s: n=%d outputs=%d" % (o.name, n, len(outs)), flush=True)
    Pt = torch.from_numpy(np.ascontiguousarray(P)).to(dev)
    out = torch.full((3 * len(outs), n), -7.0, device=dev)
    ss.execute(g, n, {"P": Pt, "N": Pt, "dPdx": dPdx, "dPdy": dPdy}, outputs=out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def inputs(n=20000, seed=4):
    rng = np.random.default_rng(seed)
    P = rng.uniform(-0.2, 1.2, (3, n)).astype(np.float32)
    P[:, :64] = np.round(P[:, :64] * 2) / 2  # greys, primaries and ties
    return P


@pytest.mark.parametrize("half", [0, 1])
def test_transformc_pairs_match_interpreter(dev, half):
    pairs = PAIRS[10 * half:10 * half + 10]
    P = inputs()
    o, outs = pair_shader("tc_pairs%d" % half, pairs)
    got = run(dev, o, outs, P)
    for k, (frm, to) in enumerate(pairs):
        ref = O.transformc(frm, to, P)
        ok = ulp_close(got[3 * k:3 * k + 3], ref)
        assert ok.all(), (frm, to, np.argwhere(~ok)[:5], got[3 * k:3 * k + 3][~ok][:5],
                          ref[~ok][:5])


def test_color_constructor_spaces(dev):
    P = inputs(seed=5)
    o, outs = pair_shader("tc_ctor", CTOR, ctor=True)
    got = run(dev, o, outs, P)
    for k, sp in enumerate(CTOR):
        assert ulp_close(got[3 * k:3 * k + 3], O.transformc(sp, None, P)).all(), sp


def deriv_shader(frm, to):
    o = G.Oso("shader", "tc_d_%s_%s" % (frm, to))
    o.sym("global", "point", "P")
    c, t = o.tmp("color"), o.tmp("color")
    for nm in ("cout", "dxo", "dyo"):
        o.sym("oparam", "color", nm, "0 0 0")
    o.op("assign", c, "P")
    o.op("transformc", t, o.const("string", frm), o.const("string", to), c)
    o.op("assign", "cout", t)
    o.op("Dx", "dxo", t)
    o.op("Dy", "dyo", t)
    return o, ["cout", "dxo", "dyo"]


GOLD = json.load(open(os.path.join(HERE, "golden", "transformc.json")))["special"]


@pytest.mark.parametrize("row", [r for r in GOLD if len(r["in"]) == 3],
                         ids=lambda r: "%s-%s" % (r["from"], r["to"]))
def test_derivatives_match_reference_golden(dev, row):
    """transformc("rgb", space, color(u, v, 0.5)) at u = v = 0.5 with
    Dx(u) = Dy(v) = 1, as testshade prints it."""
    o, outs = deriv_shader(row["from"], row["to"])
    P = np.array(row["in"][0], np.float32).reshape(3, 1)
    got = run(dev, o, outs, P, dPdx=(1.0, 0.0, 0.0), dPdy=(0.0, 1.0, 0.0))[:, 0]
    np.testing.assert_allclose(got.reshape(3, 3), np.array(row["out"]), atol=5.001e-4)


def _f64_transform(frm, to, c):
    """float64 numpy restatement of transformc for the derivative check
    (the formulas of opcolor_impl.h:246-490; matrices from the colour
    system)."""
    cs = O.colorsys()
    x2r, r2x = cs["xyz2rgb"].astype(np.float64), cs["rgb2xyz"].astype(np.float64)
    yiq2rgb = np.array([[1, 1, 1], [0.9557, -0.2716, -1.1082], [0.6199, -0.6469, 1.7051]])

    def hsv2rgb(h, s_, v):
        h6 = 6 * (h - np.floor(h))
        hi = np.floor(h6)
        f = h6 - hi
        p, q, t = v * (1 - s_), v * (1 - s_ * f), v * (1 - s_ * (1 - f))
        sel = [(v, t, p), (q, v, p), (p, v, t), (p, q, v), (t, p, v), (v, p, q)]
        out = [np.select([hi == k for k in range(5)], [sel[k][j] for k in range(5)], sel[5][j])
               for j in range(3)]
        return np.stack(out)

    def rgb2hsv(c):
        mx, mn = c.max(axis=0), c.min(axis=0)
        d = mx - mn
        s_ = np.where(mx > 0, d / np.where(mx > 0, mx, 1), 0)
        r, g, b = c
        dd = np.where(d > 0, 6 * d, 1.0)
        h = np.where(r >= mx, (g - b) / dd, np.where(g >= mx, 2 / 6 + (b - r) / dd,
                                                     4 / 6 + (r - g) / dd))
        h = np.where(s_ > 0, np.where(h < 0, h + 1, h), 0.0)
        return np.stack([h, s_, mx])

    def to_rgb(c):
        if frm == "rgb":
            return c
        if frm == "hsv":
            return hsv2rgb(*c)
        if frm == "hsl":
            h, s_, l_ = c
            v = np.where(l_ <= 0.5, l_ * (1 + s_), l_ * (1 - s_) + s_)
            return hsv2rgb(h, (v - (2 * l_ - v)) / v, v)
        if frm == "xyY":
            Y_y = c[2] / c[1]
            return (np.stack([Y_y * c[0], c[2], Y_y * (1 - c[0] - c[1])]).T @ x2r).T
        if frm == "sRGB":
            return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
        raise ValueError(frm)

    c = to_rgb(c)
    if to == "rgb":
        return c
    if to == "hsv":
        return rgb2hsv(c)
    if to == "hsl":
        hsv = rgb2hsv(c)
        mn, mx = c.min(axis=0), hsv[2]
        l_ = (mn + mx) / 2
        s_ = np.where(mn == mx, 0, (mx - mn) / np.where(l_ <= 0.5, mx + mn, 2 - mx - mn))
        return np.stack([hsv[0], s_, l_])
    if to == "xyY":
        xyz = (c.T @ r2x).T
        n = xyz.sum(axis=0)
        return np.stack([xyz[0] / n, xyz[1] / n, xyz[1]])
    if to == "sRGB":
        return np.where(c <= 0.0031308, 12.92 * c, 1.055 * c ** (1 / 2.4) - 0.055)
    if to == "YIQ":
        a = np.array([[0.299, 0.596, 0.212], [0.587, -0.275, -0.523], [0.114, -0.321, 0.311]])
        return (c.T @ a).T
    raise ValueError(to)


@pytest.mark.parametrize("pair", [("rgb", "hsv"), ("rgb", "hsl"), ("hsv", "rgb"), ("hsl", "rgb"),
                                  ("rgb", "xyY"), ("xyY", "rgb"), ("sRGB", "rgb"),
                                  ("rgb", "sRGB"), ("rgb", "YIQ")], ids="-".join)
def test_derivatives_match_central_differences(dev, pair):
    """Dual2 derivatives of transformc against central differences of a
    float64 restatement of the transform (steps 1e-6 and 2e-6; points where
    the two disagree straddle a kink of a piecewise-smooth transform). The
    inputs are drawn with no ties between components: at a tie (a grey, a
    hue-sector edge, l = 0.5) a central difference averages two one-sided
    slopes while Dual2 arithmetic takes one of them."""
    rng = np.random.default_rng(6)
    P = rng.uniform(0.05, 0.95, (3, 4000)).astype(np.float32)
    dx, dy = (1e-3, 0.0, 0.0), (0.0, 1e-3, 0.0)
    o, outs = deriv_shader(*pair)
    got = run(dev, o, outs, P, dPdx=dx, dPdy=dy)
    # the Dual2 value path is the float path
    assert ulp_close(got[0:3], O.transformc(pair[0], pair[1], P)).all()
    P64 = P.astype(np.float64)
    # the float64 restatement agrees with the oracle's values
    np.testing.assert_allclose(_f64_transform(pair[0], pair[1], P64),
                               O.transformc(pair[0], pair[1], P), rtol=2e-5, atol=2e-6)
    for axis, d in ((0, got[3:6]), (1, got[6:9])):
        est = []
        for h in (1e-6, 2e-6):
            e = np.zeros((3, 1))
            e[axis] = h
            est.append((_f64_transform(pair[0], pair[1], P64 + e)
                        - _f64_transform(pair[0], pair[1], P64 - e)) / (2 * h) * 1e-3)
        smooth = np.abs(est[0] - est[1]) <= 1e-9 + 1e-6 * np.abs(est[0])
        assert smooth.mean() > 0.95
        err = np.abs(d - est[0])[smooth]
        # binary32 Dual2 arithmetic: about 1e-6 relative, more where terms cancel
        ok = err <= 1e-8 + 1e-4 * np.abs(est[0][smooth])
        assert ok.mean() >= 0.999, (pair, axis, ok.mean(), err.max())
        assert np.median(err / np.maximum(np.abs(est[0][smooth]), 1e-9)) < 1e-5
