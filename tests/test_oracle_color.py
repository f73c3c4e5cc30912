"""The oracle's color-space transforms (ColorSystem::transformc and the
color(space, ...) constructor's to_rgb) against

  * the reference's golden output testsuite/transformc/ref/out.txt
    (hsv, hsl, YIQ, XYZ, xyY both ways, linear <-> sRGB; 3 decimals);
  * closed forms: hsv / hsl of the primaries and greys, xyY of the white
    point, the YIQ luma weights;
  * round trips rgb -> space -> rgb on random colors (the reference's own
    residual check), and the 1/12.92 and 1.055 x^(1/2.4) - 0.055 sRGB
    branches on either side of their thresholds.
A planted slip (hue sector order, hsl's l > 0.5 divisor, YIQ transposed)
fails at least one of these.
"""
import json
import os

import numpy as np
import pytest

import _oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "transformc.json")))
SPACES = ["hsv", "hsl", "YIQ", "XYZ", "xyY", "sRGB"]


def col(*v):
    return np.array(v, np.float32).reshape(3, 1)


PAIRS = [(GOLD["pairs"][k], GOLD["pairs"][k + 1]) for k in range(0, len(GOLD["pairs"]), 2)]


@pytest.mark.parametrize("fwd,back", PAIRS, ids=lambda r: r["to"] + r["from"])
def test_golden_pairs(fwd, back):
    """test_ctransform: cs = transformc(space, c); crgb = transformc(space,
    "rgb", cs), the second from the unrounded first, as the shader does."""
    assert fwd["from"] == "rgb" and back["from"] == fwd["to"] and back["to"] == "rgb"
    cs = O.transformc("rgb", fwd["to"], col(*fwd["in"]))
    np.testing.assert_allclose(cs[:, 0], fwd["out"], atol=5.001e-4)
    crgb = O.transformc(fwd["to"], "rgb", cs)
    np.testing.assert_allclose(crgb[:, 0], back["out"], atol=5.001e-4)


@pytest.mark.parametrize("row", GOLD["special"], ids=lambda r: "%s-%s" % (r["from"], r["to"]))
def test_golden_values(row):
    # the value line of each entry (derivative lines are checked on the GPU)
    got = O.transformc(row["from"], row["to"], col(*row["in"][0]))[:, 0]
    np.testing.assert_allclose(got, row["out"][0], atol=5.001e-4)


def test_closed_forms():
    prim = np.array([[1, 0, 0, 1, 0, 1, .5], [0, 1, 0, 1, 1, 0, .5], [0, 0, 1, 0, 1, 1, .5]],
                    np.float32)
    hsv = O.transformc("rgb", "hsv", prim)
    np.testing.assert_allclose(hsv[0], [0, 1 / 3, 2 / 3, 1 / 6, 1 / 2, 5 / 6, 0], atol=1e-7)
    np.testing.assert_allclose(hsv[1], [1, 1, 1, 1, 1, 1, 0])
    np.testing.assert_allclose(hsv[2], [1, 1, 1, 1, 1, 1, .5])
    hsl = O.transformc("rgb", "hsl", prim)
    np.testing.assert_allclose(hsl[0], hsv[0])
    np.testing.assert_allclose(hsl[1], [1, 1, 1, 1, 1, 1, 0])
    np.testing.assert_allclose(hsl[2], [.5, .5, .5, .5, .5, .5, .5])
    # hsl with l > 0.5: pastel red (1, .5, .5): s = (1 - .5) / (2 - 1 - .5) = 1
    np.testing.assert_allclose(O.transformc("rgb", "hsl", col(1, .5, .5))[:, 0],
                               [0, 1, .75], atol=1e-7)
    # YIQ luma is the NTSC weighting
    y = O.transformc("rgb", "YIQ", np.eye(3, dtype=np.float32))[0]
    np.testing.assert_allclose(y, [0.299, 0.587, 0.114], rtol=1e-6)
    # the white of the working space has Rec709's D65 chromaticity
    w = O.transformc("rgb", "xyY", col(1, 1, 1))[:, 0]
    np.testing.assert_allclose(w[:2], [0.3127, 0.3290], atol=2e-4)
    np.testing.assert_allclose(w[2], 1.0, atol=1e-5)


@pytest.mark.parametrize("space", SPACES)
def test_round_trip(space):
    rng = np.random.default_rng(3)
    c = rng.uniform(0.01, 1.0, (3, 20000)).astype(np.float32)
    back = O.transformc(space, "rgb", O.transformc("rgb", space, c))
    want = c.astype(np.float64)
    if space == "YIQ":
        # the 4-digit YIQ matrices are inverses only to about 1e-3: the
        # round trip is the row vector times their float64 product
        a = np.array([[0.299, 0.596, 0.212], [0.587, -0.275, -0.523], [0.114, -0.321, 0.311]])
        b = np.array([[1.0, 1.0, 1.0], [0.9557, -0.2716, -1.1082], [0.6199, -0.6469, 1.7051]])
        want = (want.T @ (a @ b)).T
    np.testing.assert_allclose(back, want, rtol=2e-4, atol=2e-5)


def test_srgb_branches():
    x = np.array([0.0, 0.003, 0.0031308, 0.004, 0.04045, 0.05, 0.5, 1.0], np.float32)
    c = np.stack([x, x, x])
    to = O.transformc("linear", "sRGB", c)[0].astype(np.float64)
    xd = x.astype(np.float64)
    want = np.where(xd <= 0.0031308, 12.92 * xd, 1.055 * xd ** (1 / 2.4) - 0.055)
    np.testing.assert_allclose(to, want, rtol=2e-6, atol=1e-7)
    back = O.transformc("sRGB", "linear", c)[0].astype(np.float64)
    want = np.where(xd <= 0.04045, xd / 12.92, ((xd + 0.055) / 1.055) ** 2.4)
    np.testing.assert_allclose(back, want, rtol=2e-6, atol=1e-7)


def test_constructor_spaces():
    """color(space, ...) converts like transformc(space, "rgb"), but knows
    neither "linear" nor "sRGB" (OpenColorIO names there)."""
    c = col(0.3, 0.6, 0.9)
    for sp in ("hsv", "hsl", "YIQ", "XYZ", "xyY"):
        np.testing.assert_array_equal(O.transformc(sp, None, c), O.transformc(sp, "rgb", c))
    for sp in ("linear", "sRGB", "ACEScg"):
        with pytest.raises(ValueError):
            O.transformc(sp, None, c)
