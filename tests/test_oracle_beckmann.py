"""Microfacet<BeckmannDist,0> evaluation in the oracle
(src/testrender/shading.cpp:510-522 with the shared Microfacet eval), pinned
by closed forms: at normal incidence (Lambda(wo) = 0) the lobe's pdf
integrates to 1 over the sphere (D's normalisation and the reflection
Jacobian), and weight = F / (1 + Lambda(wi)) recovers Lambda, which must
follow the exact Smith-Beckmann form (erf(a) - 1) / 2 + exp(-a^2) /
(2 a sqrt(pi)) within the rational approximation's error."""
import numpy as np
import pytest
from scipy.special import erf

import _oracle as O


def lobe(alpha, eta):
    return [[2, 1.0, 1.0, 1.0, 0.0, 0.0, 1.0, alpha, eta]]


def sphere(nt=720, npf=360):
    t = (np.arange(nt) + 0.5) * (np.pi / 2) / nt
    p = (np.arange(npf) + 0.5) * 2 * np.pi / npf
    T, P = np.meshgrid(t, p, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)]).reshape(3, -1)
    return d.astype(np.float32), (np.sin(T) * (np.pi / 2 / nt) * (2 * np.pi / npf)).reshape(-1)


@pytest.mark.parametrize("alpha", [0.15, 0.3])
def test_beckmann_pdf_integrates_to_one_at_normal_incidence(alpha):
    d, w = sphere()
    wo = np.repeat(np.array([[0.0], [0.0], [1.0]], np.float32), d.shape[1], 1)
    e = O.composite_eval(lobe(alpha, 1.5), wo, d)
    assert abs(float(np.sum(e[3] * w)) - 1.0) < 3e-3


def fresnel(c, eta):
    g = np.sqrt(eta * eta - 1 + c * c)
    A = (g - c) / (g + c)
    B = (c * (g + c) - 1) / (c * (g - c) + 1)
    return 0.5 * A * A * (1 + B * B)


def test_beckmann_lambda_follows_smith():
    alpha, eta = 0.3, 1.5
    th = np.linspace(0.05, 1.45, 200)
    wi = np.stack([np.sin(th), np.zeros_like(th), np.cos(th)]).astype(np.float32)
    wo = np.repeat(np.array([[0.0], [0.0], [1.0]], np.float32), th.size, 1)
    e = O.composite_eval(lobe(alpha, eta), wo, wi)
    h = wi.astype(np.float64) + wo
    m = h / np.linalg.norm(h, axis=0)
    F = fresnel(m[2], eta)            # m . wo with wo = n
    lam = F / e[0].astype(np.float64) - 1
    a = 1 / (np.tan(th) * alpha)
    exact = np.where(a < 50, (erf(a) - 1) / 2 + np.exp(-a * a) / (2 * a * np.sqrt(np.pi)), 0.0)
    assert np.max(np.abs(lam - exact)) < 4e-3
