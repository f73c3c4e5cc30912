"""Pin the oracle's testrender closures (oracle/closure.cpp) and its OSO
interpreter (oracle/osoint.cpp) to what mathematics and an independent
restatement fix.

Closures (the reference tree holds no numeric fixtures for testrender's
lobes, so these are closed forms and quadratures):
  * Diffuse: pdf = cos/pi and weight 1 (cosine-weighted sampling);
  * GGX reflection: the reflected-VNDF pdf integrates to 1 over the sphere
    for a near-normal view and low roughness; with a Fresnel of 1 the
    directional albedo (integral of weight * pdf) is close to 1 at low
    roughness and falls with roughness (single scattering);
  * the GGX Smith Lambda against the textbook form (-1 + sqrt(1 + a^2
    tan^2))/2, through the lobe's own weight G2/G1;
  * CompositeBSDF's MIS combination: two diffuse lobes of weights a and b
    give weight a + b and pdf cos/pi.
Interpreter: the fbm_cell shader (OSO) run by the interpreter equals the
oracle's hand restatement of the same shader (oracle/shade.cpp), which
shares no control flow with it.
"""
import os

import numpy as np
import pytest

import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _sphere_grid(nt=400, np_=800):
    """Midpoint quadrature nodes over the sphere and their solid angles."""
    t = (np.arange(nt) + 0.5) * np.pi / nt
    p = (np.arange(np_) + 0.5) * 2 * np.pi / np_
    T, Pp = np.meshgrid(t, p, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)]).reshape(3, -1)
    dw = (np.sin(T) * (np.pi / nt) * (2 * np.pi / np_)).ravel()
    return d, dw


def _lobe(kind, w=(1, 1, 1), N=(0, 0, 1), alpha=0.0, eta=0.0):
    return [kind, *w, *N, alpha, eta]


def test_diffuse_closed_form():
    rng = np.random.default_rng(0)
    wi = rng.normal(size=(3, 1000))
    wi /= np.linalg.norm(wi, axis=0)
    wo = np.tile(np.array([[0.0], [0.0], [1.0]]), (1, 1000))
    out = O.composite_eval([_lobe(0, w=(0.3, 0.5, 0.7))], wo, wi)
    cos = np.maximum(wi[2], 0)
    np.testing.assert_allclose(out[3], cos / np.pi, rtol=1e-6, atol=1e-7)
    up = cos > 0
    np.testing.assert_allclose(out[0][up], 0.3, rtol=1e-6)
    np.testing.assert_allclose(out[2][up], 0.7, rtol=1e-6)


def _vndf_above_horizon(alpha, wo, ns=2000, nphi=720):
    """Fraction of GGX visible normals (textbook D and Smith G1, float64,
    quadrature over microfacet normals) whose mirror reflection of wo lies
    above the surface."""
    s = (np.arange(ns) + 0.5) * (np.pi / 2) / ns
    th = np.arctan(alpha * np.tan(s))  # concentrates nodes near the peak
    dth = alpha / np.cos(s) ** 2 / (1 + (alpha * np.tan(s)) ** 2) * (np.pi / 2) / ns
    ph = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    m = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)])
    dm = (np.sin(T) * dth[:, None]) * (2 * np.pi / nphi)
    tan2 = np.tan(T) ** 2
    D = alpha ** 2 / (np.pi * np.cos(T) ** 4 * (alpha ** 2 + tan2) ** 2)
    tan2o = (1 - wo[2] ** 2) / wo[2] ** 2
    G1 = 1 / (1 + (-1 + np.sqrt(1 + alpha ** 2 * tan2o)) / 2)
    om = np.einsum("i,ijk->jk", wo, m)
    dv = G1 * np.maximum(om, 0) * D / wo[2]
    wi_z = 2 * om * m[2] - wo[2]
    return float((dv * dm * (wi_z > 0)).sum()), float((dv * dm).sum())


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5])
def test_ggx_pdf_matches_visible_normal_measure(alpha):
    """The lobe's pdf over reflected directions integrates to the visible-
    normal measure of the normals that reflect above the surface (the
    Jacobian 1/(4 wo.m) and G1 D / wo.z together)."""
    d, dw = _sphere_grid(800, 1200)
    th = 0.2
    wov = np.array([np.sin(th), 0.0, np.cos(th)])
    wo = np.tile(wov[:, None], (1, d.shape[1]))
    out = O.composite_eval([_lobe(1, alpha=alpha, eta=1.5)], wo, d)
    total = float((out[3].astype(np.float64) * dw).sum())
    above, whole = _vndf_above_horizon(alpha, wov)
    assert abs(whole - 1.0) < 2e-3  # the quadrature itself is normalised
    assert abs(total - above) < 3e-3, (total, above)


def test_ggx_white_furnace():
    """eta = 0 makes testrender's Fresnel 1; the single-scattering albedo
    (integral of weight * pdf = f cos) then equals a float64 quadrature of
    the textbook Cook-Torrance integrand D G2 / (4 cos_o) (half-vector
    form, height-correlated Smith), tends to 1 at low roughness and falls
    as roughness grows."""
    d, dw = _sphere_grid()
    up = d[2] > 0
    alb = []
    for alpha, th in ((0.05, 0.0), (0.3, 0.7), (0.8, 0.3)):
        wov = np.array([np.sin(th), 0.0, np.cos(th)])
        wo = np.tile(wov[:, None], (1, d.shape[1]))
        out = O.composite_eval([_lobe(1, alpha=alpha, eta=0.0)], wo, d)
        got = float((out[0].astype(np.float64) * out[3] * dw).sum())
        h = d + wov[:, None]
        h /= np.linalg.norm(h, axis=0)
        tan2h = (1 - h[2] ** 2) / h[2] ** 2
        D = alpha ** 2 / (np.pi * h[2] ** 4 * (alpha ** 2 + tan2h) ** 2)

        def lam(c):
            return (-1 + np.sqrt(1 + alpha ** 2 * (1 - c * c) / np.maximum(c * c, 1e-300))) / 2
        G2 = 1 / (1 + lam(wov[2]) + lam(d[2]))
        ref = float((D * G2 / (4 * wov[2]) * dw * up).sum())
        assert abs(got - ref) < 1e-5 * ref + 1e-7, (alpha, got, ref)
        alb.append(got)
    assert abs(alb[0] - 1.0) < 5e-3
    assert alb[0] > alb[1] > alb[2] > 0.3


def test_ggx_weight_is_smith_ratio():
    """weight = F * G2/G1 with G1 = 1/(1+Lo), G2 = 1/(1+Lo+Li) and the
    textbook Lambda; F = 1 through eta = 0."""
    rng = np.random.default_rng(1)
    n = 2000
    alpha = 0.35

    def hemi():
        v = rng.normal(size=(3, n))
        v[2] = np.abs(v[2]) + 0.05
        return v / np.linalg.norm(v, axis=0)
    wo, wi = hemi(), hemi()
    out = O.composite_eval([_lobe(1, alpha=alpha, eta=0.0)], wo, wi)

    def lam(w):
        tan2 = (1 - w[2] ** 2) / w[2] ** 2
        return (-1 + np.sqrt(1 + alpha * alpha * tan2)) / 2
    ref = (1 + lam(wo)) / (1 + lam(wo) + lam(wi))
    np.testing.assert_allclose(out[0], ref, rtol=2e-5)


def test_composite_mis_of_two_diffuse_lobes():
    rng = np.random.default_rng(2)
    wi = rng.normal(size=(3, 500))
    wi[2] = np.abs(wi[2]) + 0.01
    wi /= np.linalg.norm(wi, axis=0)
    wo = np.tile(np.array([[0.0], [0.0], [1.0]]), (1, 500))
    out = O.composite_eval([_lobe(0, w=(0.2, 0.2, 0.2)), _lobe(0, w=(0.6, 0.6, 0.6))], wo, wi)
    np.testing.assert_allclose(out[0], 0.8, rtol=1e-6)
    np.testing.assert_allclose(out[3], wi[2] / np.pi, rtol=1e-6)


def test_interpreter_matches_restated_fbm_cell():
    with open(os.path.join(ROOT, "openshadinglanguage_amd", "shaders", "fbm_cell.oso")) as f:
        text = f.read()
    g = O.Group([("fbm_cell", "fbm", text)], []).outputs("", "Cout")
    rng = np.random.default_rng(3)
    P = rng.uniform(-50, 150, size=(3, 4000)).astype(np.float32)
    N = np.zeros_like(P)
    N[2] = 1
    out, _ = g.run(P, N)
    ref = O.shade_c1(P)
    np.testing.assert_array_equal(out[1:4], ref)
