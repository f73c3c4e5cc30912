"""testrender's Phong and Diffuse<1> (translucent) lobes in the oracle
(src/testrender/shading.cpp:300-362), pinned by:
- Phong at normal incidence (wo = N, so the mirror direction is N): the
  pdf is the normalised cosine-power lobe, integrating to 1 over the
  hemisphere (float64 quadrature), and the weight is cos(N, wi) (e + 2) /
  (e + 1);
- sampled directions carry the eval's weight and pdf, and the mean sampled
  weight equals the directional reflectance by quadrature;
- the translucent lobe is the diffuse lobe about -N."""
import numpy as np
import pytest

import _oracle as O


def _sphere_grid(nt=600, nph=1200):
    th = (np.arange(nt) + 0.5) * np.pi / nt
    ph = (np.arange(nph) + 0.5) * 2 * np.pi / nph
    T, Ph = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(Ph), np.sin(T) * np.sin(Ph), np.cos(T)]).reshape(3, -1)
    dw = (np.sin(T) * (np.pi / nt) * (2 * np.pi / nph)).reshape(-1)
    return d, dw


@pytest.mark.parametrize("e", [2.0, 10.0, 45.5])
def test_phong_normal_incidence(e):
    N = np.array([0.0, 0.0, 1.0])
    wi, dw = _sphere_grid()
    wo = np.repeat(N[:, None], wi.shape[1], 1)
    r = O.composite_eval(np.array([[5, 1, 1, 1, *N, e, 0.0]]), wo, wi).astype(np.float64)
    assert abs((r[3] * dw).sum() - 1.0) < 2e-3
    # where the lobe's pdf underflows the composite no longer reports the
    # lobe's own weight: compare where it is representable
    up = (wi[2] > 1e-3) & (r[3] > 1e-6)
    np.testing.assert_allclose(r[0, up], wi[2, up] * (e + 2) / (e + 1), rtol=2e-6)
    assert np.all(r[:, wi[2] <= 0] == 0)


@pytest.mark.parametrize("e", [3.0, 20.0])
def test_phong_samples_agree_with_eval_and_reflectance(e):
    rng = np.random.default_rng(int(e))
    N = np.array([0.0, 0.0, 1.0])
    lobe = np.array([[5, 1, 1, 1, *N, e, 0.0]])
    wo1 = np.array([[np.sin(0.6)], [0.0], [np.cos(0.6)]])
    n = 200000
    wo = np.repeat(wo1, n, 1)
    s = O.composite_sample(lobe, wo, rng.uniform(size=(3, n)))
    ok = s[6] > 1e-3
    assert ok.mean() > 0.5
    ev = O.composite_eval(lobe, wo[:, ok][:, :5000], s[:3, ok][:, :5000])
    np.testing.assert_allclose(ev[3], s[6, ok][:5000], rtol=1e-5)
    np.testing.assert_allclose(ev[0], s[3, ok][:5000], rtol=1e-5)
    mc = float(np.where(ok, s[3], 0.0).mean())
    wi, dw = _sphere_grid()
    q = O.composite_eval(lobe, np.repeat(wo1, wi.shape[1], 1), wi).astype(np.float64)
    quad = float((q[0] * q[3] * dw).sum())
    assert abs(mc - quad) < 0.01 * quad + 3 / np.sqrt(n), (mc, quad)


def test_translucent_is_diffuse_about_minus_n():
    rng = np.random.default_rng(2)
    n = 3000
    N = np.array([0.2, -0.3, 0.9]) / np.linalg.norm([0.2, -0.3, 0.9])
    wo = rng.normal(size=(3, n))
    wi = rng.normal(size=(3, n))
    wi /= np.linalg.norm(wi, axis=0)
    tr = O.composite_eval(np.array([[4, 1, 1, 1, *(-N), 0.0, 0.0]]), wo, wi)
    np.testing.assert_allclose(tr[3], np.maximum(-(N @ wi), 0) / np.pi, rtol=1e-5, atol=1e-7)
    assert np.all(tr[0][tr[3] > 0] == 1.0)


def test_delta_lobes():
    """reflection / refraction / transparent (shading.cpp:1081-1149): the
    mirror direction with the dielectric Fresnel weight (a perfect mirror
    without eta), the Snell direction with the transmitted share, -wo; all
    with an infinite pdf and nothing to evaluate. Reflected and refracted
    weights of one interface sum to 1."""
    rng = np.random.default_rng(4)
    n = 2000
    N = np.array([0.0, 0.0, 1.0])
    wo = rng.normal(size=(3, n))
    wo[2] = np.abs(wo[2]) + 0.05
    wo /= np.linalg.norm(wo, axis=0)
    rnd = rng.uniform(size=(3, n))
    eta = 1.5
    refl = O.composite_sample(np.array([[7, 1, 1, 1, *N, 0.0, eta]]), wo, rnd)
    mirror = O.composite_sample(np.array([[6, 1, 1, 1, *N, 0.0, 0.0]]), wo, rnd)
    refr = O.composite_sample(np.array([[8, 1, 1, 1, *N, 0.0, eta]]), wo, rnd)
    tr = O.composite_sample(np.array([[9, 1, 1, 1, 0, 0, 0, 0.0, 0.0]]), wo, rnd)
    want_r = 2 * wo[2] * N[:, None] - wo
    np.testing.assert_allclose(refl[:3], want_r, atol=2e-7)
    c = wo[2].astype(np.float64)
    g = np.sqrt(eta * eta - 1 + c * c)
    A, B = (g - c) / (g + c), (c * (g + c) - 1) / (c * (g - c) + 1)
    F = 0.5 * A * A * (1 + B * B)
    np.testing.assert_allclose(refl[3], F, rtol=2e-6)
    np.testing.assert_allclose(mirror[3], 1.0)
    np.testing.assert_allclose(refl[3] + refr[3], 1.0, rtol=2e-6)
    # Snell: sin(theta_t) = sin(theta_i) / eta, below the surface, in the plane
    sin_i = np.sqrt(1 - c * c)
    sin_t = np.sqrt(1 - refr[2].astype(np.float64) ** 2)
    np.testing.assert_allclose(sin_t, sin_i / eta, atol=2e-6)
    assert np.all(refr[2] < 0)
    np.testing.assert_allclose(tr[:3], -wo, atol=0)
    for s in (refl, mirror, refr, tr):
        assert np.all(np.isinf(s[6]))
    ev = O.composite_eval(np.array([[7, 1, 1, 1, *N, 0.0, eta]]), wo, want_r)
    assert np.all(ev == 0)
