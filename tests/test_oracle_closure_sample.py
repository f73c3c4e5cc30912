"""CompositeBSDF::sample of the oracle (Diffuse, GGX microfacet and their
MIS combination; src/testrender/shading.h:376-402, shading.cpp:313-320,
:655-685), pinned by (1) agreement of each lobe's sampled weight and pdf
with the independently pinned eval at the drawn direction, and (2) the
Monte Carlo identity E[weight over directions above the surface] =
integral of the BSDF times cosine, the integral taken by quadrature of
eval."""
import numpy as np
import pytest

import _oracle as O


def unit(v):
    return v / np.linalg.norm(v, axis=0, keepdims=True)


def lobe(kind, w, N, alpha=0.0, eta=0.0):
    return [kind, w[0], w[1], w[2], N[0], N[1], N[2], alpha, eta]


N0 = (0.0, 0.0, 1.0)
LOBES = {
    "diffuse": [lobe(0, (1, 1, 1), N0)],
    "ggx": [lobe(1, (1, 1, 1), N0, 0.3, 1.5)],
    "ggx_rough": [lobe(1, (1, 1, 1), N0, 0.8, 2.2)],
    "mix": [lobe(0, (0.6, 0.3, 0.2), N0), lobe(1, (1, 1, 1), N0, 0.25, 1.6)],
}


def wo_set(n, rng, min_cos=0.15):
    th = np.arccos(rng.uniform(min_cos, 1.0, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]).astype(np.float32)


@pytest.mark.parametrize("name", ["diffuse", "ggx", "ggx_rough"])
def test_single_lobe_sample_agrees_with_eval(name):
    rng = np.random.default_rng(1)
    n = 20000
    wo = wo_set(n, rng)
    rnd = rng.uniform(0, 1, (3, n)).astype(np.float32)
    s = O.composite_sample(LOBES[name], wo, rnd)
    # the microfacet sampler, like testrender's, also returns reflections
    # that end below the surface (eval is zero there): compare the rest
    ok = (s[6] > 0) & (s[2] > 0)
    assert ok.mean() > 0.6
    wi = s[0:3, ok]
    e = O.composite_eval(LOBES[name], wo[:, ok], wi)
    np.testing.assert_allclose(s[6, ok], e[3], rtol=2e-4, atol=1e-6)
    np.testing.assert_allclose(s[3:6, ok], e[0:3], rtol=2e-4, atol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(wi, axis=0), 1.0, atol=2e-6)


def sphere_quadrature(nt=360, npf=720):
    t = (np.arange(nt) + 0.5) * np.pi / nt
    p = (np.arange(npf) + 0.5) * 2 * np.pi / npf
    T, P = np.meshgrid(t, p, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)]).reshape(3, -1)
    w = (np.sin(T) * (np.pi / nt) * (2 * np.pi / npf)).reshape(-1)
    return d.astype(np.float32), w


@pytest.mark.parametrize("name", ["diffuse", "ggx", "mix"])
def test_sample_mean_weight_is_reflectance(name):
    rng = np.random.default_rng(2)
    lobes = LOBES[name]
    d, w = sphere_quadrature()
    for cos_o in (0.9, 0.5):
        wo1 = np.array([np.sqrt(1 - cos_o ** 2), 0.0, cos_o], np.float32)
        # reflectance: sum over lobes of lobe weight x (eval weight x pdf)
        refl = np.zeros(3)
        for lb in lobes:
            single = [lb[:1] + [1.0, 1.0, 1.0] + lb[4:]]
            e = O.composite_eval(single, np.repeat(wo1[:, None], d.shape[1], 1), d)
            refl += np.asarray(lb[1:4]) * np.sum(e[0] * e[3] * w)
        n = 400000
        rnd = rng.uniform(0, 1, (3, n)).astype(np.float32)
        s = O.composite_sample(lobes, np.repeat(wo1[:, None], n, 1), rnd)
        # directions below the surface carry no reflectance (eval is zero
        # there) but the sampler, like testrender's, can return them
        wgt = s[3:6].astype(np.float64) * (s[2] > 0)
        est = wgt.mean(axis=1)
        sd = wgt.std(axis=1) / np.sqrt(n)
        assert np.all(np.abs(est - refl) <= 5 * sd + 2e-3 * refl), (name, cos_o, est, refl, sd)


def test_diffuse_sampling_is_cosine_weighted():
    rng = np.random.default_rng(3)
    n = 200000
    lobes = [lobe(0, (1, 1, 1), (0.6, 0.0, 0.8))]
    wo = np.repeat(np.array([[0.0], [0.0], [1.0]], np.float32), n, 1)
    s = O.composite_sample(lobes, wo, rng.uniform(0, 1, (3, n)).astype(np.float32))
    c = s[0] * 0.6 + s[2] * 0.8  # cosine to N
    assert c.min() >= -1e-6
    assert abs(c.mean() - 2.0 / 3.0) < 3e-3      # E[cos] under cos/pi
    assert abs((c ** 2).mean() - 0.5) < 3e-3     # E[cos^2]
    np.testing.assert_allclose(s[6], c / np.pi, rtol=1e-4, atol=1e-6)
