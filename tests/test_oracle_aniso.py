"""Anisotropic microfacet lobes of the oracle (testrender's Microfacet<GGX|
Beckmann, 0> with xalpha != yalpha in the frame from_normal_and_tangent(N,
U), src/testrender/shading.cpp:570-800, sampling.h:30-42), pinned by:
- float64 closed forms at normal incidence (wo = N: G1 = 1, pdf = D / 4,
  weight = G2 Fr) with the frame built independently: a swapped pair of
  roughnesses, a tangent taken about the wrong axis or a dropped term fails;
- symmetry: (U, ax, ay) and (N x U, ay, ax) describe the same lobe;
- equal roughnesses: any tangent gives the isotropic lobe;
- sampling: a drawn direction's weight and pdf equal the eval there."""
import numpy as np
import pytest

import _oracle as O


def _unit(v):
    return v / np.linalg.norm(v, axis=0)


def _frame(N, U):
    u = U - N * (N @ U)
    u = u / np.linalg.norm(u)
    return u, np.cross(N, u)


def _fresnel(c, eta):
    g = np.sqrt(eta * eta - 1 + c * c)
    A = (g - c) / (g + c)
    B = (c * (g + c) - 1) / (c * (g - c) + 1)
    return 0.5 * A * A * (1 + B * B)


def _hemisphere(n, N, rng):
    w = _unit(rng.normal(size=(3, n)))
    w[:, (N @ w) < 0] *= -1
    return w


@pytest.mark.parametrize("kind", [1, 2])  # GGX, Beckmann
@pytest.mark.parametrize("ax,ay", [(0.15, 0.45), (0.6, 0.2), (0.3, 0.3)])
def test_normal_incidence_closed_form(kind, ax, ay):
    rng = np.random.default_rng(kind * 10 + int(ax * 100))
    n, eta = 3000, 1.5
    N = _unit(np.array([0.3, -0.2, 1.0]))
    U = np.array([1.0, 0.4, 0.1])
    wi = _hemisphere(n, N, rng)
    wi = wi[:, (N @ wi) > 0.05]  # away from grazing, where float32 rounding grows
    wo = np.repeat(N[:, None], wi.shape[1], 1)
    got = O.composite_eval(np.array([[kind, 1, 1, 1, *N, ax, eta, *U, ay]]), wo, wi)
    u, v = _frame(N, U)
    wl = np.stack([u @ wi, v @ wi, N @ wi])
    m = _unit(wl + np.array([[0.0], [0.0], [1.0]]))
    t2 = (m[0] ** 2 / ax ** 2 + m[1] ** 2 / ay ** 2) / m[2] ** 2  # tan^2 of m, stretched
    if kind == 1:
        D = 1 / (np.pi * ax * ay * m[2] ** 4 * (1 + t2) ** 2)
        lam = 0.5 * (-1 + np.sqrt(1 + (wl[0] ** 2 * ax ** 2 + wl[1] ** 2 * ay ** 2) / wl[2] ** 2))
    else:
        D = np.exp(-t2) / (np.pi * ax * ay * m[2] ** 4)
        a = wl[2] / np.sqrt(wl[0] ** 2 * ax ** 2 + wl[1] ** 2 * ay ** 2)
        lam = np.where(a < 1.6, (1 - 1.259 * a + 0.396 * a * a) / (3.535 * a + 2.181 * a * a), 0.0)
    want_w = _fresnel(m[2], eta) / (1 + lam)
    np.testing.assert_allclose(got[3], D / 4, rtol=2e-5 if kind == 1 else 2e-4)
    np.testing.assert_allclose(got[0], want_w, rtol=2e-4)
    # the roughnesses belong to the tangent's axis: swapping them is far off
    swapped = O.composite_eval(np.array([[kind, 1, 1, 1, *N, ay, eta, *U, ax]]), wo, wi)
    if ax != ay:
        assert np.median(np.abs(swapped[3] - got[3]) / got[3]) > 0.05


@pytest.mark.parametrize("kind", [1, 2])
def test_swapping_tangent_and_roughness_is_the_same_lobe(kind):
    rng = np.random.default_rng(7 + kind)
    n = 3000
    N = _unit(np.array([-0.1, 0.5, 1.0]))
    U = np.array([0.2, 1.0, -0.3])
    V = np.cross(N, _frame(N, U)[0])
    wo, wi = _hemisphere(n, N, rng), _hemisphere(n, N, rng)
    a = O.composite_eval(np.array([[kind, 1, 1, 1, *N, 0.2, 1.4, *U, 0.5]]), wo, wi)
    b = O.composite_eval(np.array([[kind, 1, 1, 1, *N, 0.5, 1.4, *V, 0.2]]), wo, wi)
    ok = (a[3] > 1e-3)
    np.testing.assert_allclose(b[:, ok], a[:, ok], rtol=5e-4)


def test_equal_roughness_with_any_tangent_is_isotropic():
    rng = np.random.default_rng(3)
    n = 3000
    N = _unit(np.array([0.4, 0.1, 1.0]))
    wo, wi = _hemisphere(n, N, rng), _hemisphere(n, N, rng)
    iso = O.composite_eval(np.array([[1, 1, 1, 1, *N, 0.3, 1.5]]), wo, wi)
    for U in ([1.0, 0.0, 0.0], [0.3, -0.8, 0.2], [0.0, 0.0, 0.0]):
        an = O.composite_eval(np.array([[1, 1, 1, 1, *N, 0.3, 1.5, *U, 0.3]]), wo, wi)
        ok = iso[3] > 1e-3
        np.testing.assert_allclose(an[:, ok], iso[:, ok], rtol=5e-4)
    # a zero tangent is the isotropic frame itself: bit for bit
    zero = O.composite_eval(np.array([[1, 1, 1, 1, *N, 0.3, 1.5, 0.0, 0.0, 0.0, 0.3]]), wo, wi)
    np.testing.assert_array_equal(zero, iso)


def test_sampled_directions_agree_with_eval():
    rng = np.random.default_rng(11)
    n = 4000
    N = _unit(np.array([0.2, 0.3, 1.0]))
    lobe = np.array([[1, 1, 1, 1, *N, 0.1, 1.5, 1.0, 0.2, 0.0, 0.6]])
    wo = _hemisphere(n, N, rng)
    s = O.composite_sample(lobe, wo, rng.uniform(size=(3, n)))
    # testrender's sample() keeps reflections that end below the surface,
    # which eval() gives zero: compare the ones above it
    ok = (s[6] > 1e-2) & ((N @ s[:3]) > 1e-3)
    assert ok.mean() > 0.7
    ev = O.composite_eval(lobe, wo[:, ok], s[:3, ok])
    np.testing.assert_allclose(ev[3], s[6, ok], rtol=1e-3)
    np.testing.assert_allclose(ev[0], s[3, ok], rtol=1e-3)


# --- refraction (Microfacet<GGX, 1> and <GGX, 2>) ---------------------------

def _grid_lower(N, nt=500, nph=1000):
    """Directions of the lower hemisphere about N on a (theta, phi) grid
    with their solid-angle weights."""
    u, v = _frame(N, np.array([1.0, 0.0, 0.0]))
    th = (np.arange(nt) + 0.5) * (np.pi / 2) / nt
    ph = (np.arange(nph) + 0.5) * 2 * np.pi / nph
    T, Ph = np.meshgrid(th, ph, indexing="ij")
    d = (np.sin(T) * np.cos(Ph))[None] * u[:, None, None] + (np.sin(T) * np.sin(Ph))[None] * \
        v[:, None, None] - np.cos(T)[None] * N[:, None, None]
    dw = np.sin(T) * (np.pi / 2 / nt) * (2 * np.pi / nph)
    return d.reshape(3, -1), dw.reshape(-1)


def test_transmission_pdf_is_normalised():
    """The transmission pdf is the visible-normal density carried through
    the refraction's Jacobian eta^2 |wi.m| / |ht|^2: over the lower
    hemisphere it integrates to the share of visible normals that refract
    below the surface, just under 1 for a moderate roughness at 30 degrees
    into glass. Dropping the eta^2 or the 1/|ht|^2 moves the integral by
    the factor eta^2."""
    N = np.array([0.0, 0.0, 1.0])
    wo = np.array([[np.sin(0.5)], [0.0], [np.cos(0.5)]])
    wi, dw = _grid_lower(N)
    lobe = np.array([[1, 1, 1, 1, *N, 0.3, 1.5, 0, 0, 0, 0.3, 1]])
    ev = O.composite_eval(lobe, np.repeat(wo, wi.shape[1], 1), wi)
    total = float((ev[3].astype(np.float64) * dw).sum())
    assert 0.95 < total < 1.001, total


def test_refracted_samples_agree_with_eval():
    """Refract = 1: every sample refracts; its weight and pdf equal the eval
    at the drawn direction (sampling goes through fresnel_refraction of the
    sampled microfacet, the eval through the refraction's half vector).
    Refract = 2 with the same random numbers draws the same micronormal:
    where it refracts it returns the same direction with the pdf scaled by
    the transmittance Ft and the weight divided by it; elsewhere it
    reflects, in a share of samples near the Fresnel reflectance."""
    rng = np.random.default_rng(21)
    n = 6000
    N = _unit(np.array([0.1, -0.2, 1.0]))
    row = [1, 1, 1, 1, *N, 0.2, 1.45, 1.0, 0.0, 0.0, 0.35]
    wo = _hemisphere(n, N, rng)
    rnd = rng.uniform(size=(3, n))
    s1 = O.composite_sample(np.array([row + [1]]), wo, rnd)
    below = (N @ s1[:3]) < -1e-3
    assert below.mean() > 0.8
    ok = below & (s1[6] > 1e-2)
    ev = O.composite_eval(np.array([row + [1]]), wo[:, ok], s1[:3, ok])
    np.testing.assert_allclose(ev[3], s1[6, ok], rtol=2e-3)
    np.testing.assert_allclose(ev[0], s1[3, ok], rtol=2e-3)
    s2 = O.composite_sample(np.array([row + [2]]), wo, rnd)
    same = np.all(s2[:3] == s1[:3], axis=0) & (s1[6] > 0)
    refl = ~same
    assert 0.02 < refl.mean() < 0.3
    ft = s2[6, same] / s1[6, same]
    assert np.all((ft > 0) & (ft <= 1.0000001))
    np.testing.assert_allclose(s2[3, same] * ft, s1[3, same], rtol=1e-5)
    ev2 = O.composite_eval(np.array([row + [2]]), wo[:, ok & same], s2[:3, ok & same])
    np.testing.assert_allclose(ev2[3], s2[6, ok & same], rtol=2e-3)


def test_both_splits_reflection_by_fresnel():
    """Refract = 2 reflects with weight G2/G1 and pdf x Fr where Refract = 0
    has weight G2/G1 x Fr: the ratio is the dielectric Fresnel term of the
    half vector, computed here in float64."""
    rng = np.random.default_rng(9)
    n = 3000
    N = np.array([0.0, 0.0, 1.0])
    wo, wi = _hemisphere(n, N, rng), _hemisphere(n, N, rng)
    eta = 1.5
    r0 = O.composite_eval(np.array([[1, 1, 1, 1, *N, 0.3, eta, 0, 0, 0, 0.3, 0]]), wo, wi)
    r2 = O.composite_eval(np.array([[1, 1, 1, 1, *N, 0.3, eta, 0, 0, 0, 0.3, 2]]), wo, wi)
    m = _unit(wo + wi)
    fr = _fresnel((m * wo).sum(0), eta)
    ok = r0[3] > 1e-3
    np.testing.assert_allclose(r0[0, ok] / r2[0, ok], fr[ok], rtol=1e-4)
    np.testing.assert_allclose(r2[3, ok] / r0[3, ok], fr[ok], rtol=1e-4)


# --- Beckmann sampling ---------------------------------------------------------

def test_erf_and_inverse_erf_accuracy():
    """The oracle's erf (Abramowitz-Stegun 7.1.28) and erf^-1 (Giles 2011)
    that Beckmann slope sampling uses in place of OpenImageIO's fast_erf /
    fast_ierf, against float64: erf within 2e-6 absolute (the method's 3e-7
    plus binary32 rounding of the base amplified by the 16th power), erf^-1
    within 2e-6 relative on (-0.999, 0.999) and 5e-5 at the tails, ln within
    2 ulp."""
    import math
    x = np.linspace(-4, 4, 20001).astype(np.float32)
    want = np.array([math.erf(float(v)) for v in x])
    assert np.max(np.abs(O.omath("erf", x) - want)) < 2e-6
    y = np.linspace(-0.999, 0.999, 20001).astype(np.float32)
    inv = O.omath("ierf", y).astype(np.float64)
    back = np.array([math.erf(float(v)) for v in inv])
    # d erf / dx = 2/sqrt(pi) exp(-x^2): turn the residual into an x error
    slope = 2 / np.sqrt(np.pi) * np.exp(-inv ** 2)
    assert np.max(np.abs((back - y) / slope) / np.maximum(np.abs(inv), 1e-3)) < 2e-6
    t = np.array([0.9999, 0.99999, -0.99999], np.float32)
    ti = O.omath("ierf", t).astype(np.float64)
    tb = np.array([math.erf(float(v)) for v in ti])
    assert np.all(np.abs(tb - t) / (2 / np.sqrt(np.pi) * np.exp(-ti ** 2)) < 5e-5 * np.abs(ti))
    assert np.isinf(O.omath("ierf", np.array([1.0, -1.0], np.float32))).all()
    z = np.exp(np.linspace(-80, 80, 10001)).astype(np.float32)
    lz = O.omath("log", z).astype(np.float64)
    ulp = np.spacing(np.abs(np.log(z.astype(np.float64))).astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(lz - np.log(z.astype(np.float64))) / np.maximum(ulp, 1e-30)) <= 2.0


@pytest.mark.parametrize("ax,ay", [(0.2, 0.2), (0.1, 0.4)])
def test_beckmann_samples_agree_with_eval_and_the_reflectance(ax, ay):
    """Sampled Beckmann directions (visible slopes) carry the weight and pdf
    the eval gives there, and the mean sampled weight equals the lobe's
    directional reflectance, the integral of eval weight x pdf over the
    upper hemisphere (float64 quadrature)."""
    rng = np.random.default_rng(31)
    N = np.array([0.0, 0.0, 1.0])
    lobe = np.array([[2, 1, 1, 1, *N, ax, 1.5, 1.0, 0.0, 0.0, ay]])
    n = 200000
    wo1 = np.array([[np.sin(0.7)], [0.3 * np.sin(0.7)], [np.cos(0.7)]])
    wo1 = wo1 / np.linalg.norm(wo1)
    wo = np.repeat(wo1, n, 1)
    s = O.composite_sample(lobe, wo, rng.uniform(size=(3, n)))
    up = (N @ s[:3]) > 1e-3
    ok = up & (s[6] > 1e-2)
    ev = O.composite_eval(lobe, wo[:, ok][:, :5000], s[:3, ok][:, :5000])
    np.testing.assert_allclose(ev[3], s[6, ok][:5000], rtol=2e-3)
    np.testing.assert_allclose(ev[0], s[3, ok][:5000], rtol=2e-3)
    mc = float(np.where(up, s[3], 0.0).mean())
    th = (np.arange(400) + 0.5) * (np.pi / 2) / 400
    ph = (np.arange(800) + 0.5) * 2 * np.pi / 800
    T, Ph = np.meshgrid(th, ph, indexing="ij")
    wi = np.stack([np.sin(T) * np.cos(Ph), np.sin(T) * np.sin(Ph), np.cos(T)]).reshape(3, -1)
    dw = (np.sin(T) * (np.pi / 2 / 400) * (2 * np.pi / 800)).reshape(-1)
    q = O.composite_eval(lobe, np.repeat(wo1, wi.shape[1], 1), wi).astype(np.float64)
    quad = float((q[0] * q[3] * dw).sum())
    assert abs(mc - quad) < 0.01 * quad + 3 / np.sqrt(n), (mc, quad)
