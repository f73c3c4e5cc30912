"""The oracle's binary32 Gabor noise against its float64 restatement
(tests/_gabor64.py) in every mode the batched path has: isotropic,
anisotropic and hybrid orientations, filtered and unfiltered, 1-3D, along
the axis and off-axis directions, and the layered3 group's configuration
(2D, anisotropic along x, bandwidth 1, 16 impulses, filtered by the
screen derivatives of Q = 8 P with dPdx = dPdy = 8/1024 = 0.0078125).

The isotropic image test pins the impulse stream (fast_rng, Poisson count,
impulse order); this pins what the impulses contribute: the kernel's
envelope, harmonic, direction and phase, the 2D slice and the filter
convolution. Derivatives of the oracle's Dual2 result are checked against
central differences of the float64 restatement with the filter held fixed.
"""
import numpy as np
import pytest

import _gabor64 as G64
import _oracle as O


def _points(dim, n, seed, dscale):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-6, 6, (dim, n)).astype(np.float32)
    dx = np.zeros((dim, n), np.float32)
    dy = np.zeros((dim, n), np.float32)
    dx[0] = rng.uniform(0.5, 1.0, n) * dscale
    if dim > 1:
        dy[1] = rng.uniform(0.5, 1.0, n) * dscale
        dx[1] = rng.uniform(-0.3, 0.3, n) * dscale
    if dim > 2:
        dx[2] = rng.uniform(-0.5, 0.5, n) * dscale
        dy[2] = rng.uniform(-0.5, 0.5, n) * dscale
    return p, dx, dy


def _vec3(a, i):
    v = np.zeros(3)
    v[:a.shape[0]] = a[:, i]
    return v


def _f64(p, dx, dy, i, **kw):
    return G64.gabor(_vec3(p, i), _vec3(dx, i), _vec3(dy, i), **kw)


MODES = [(dim, aniso, filt, d) for dim in (1, 2, 3) for aniso in (0, 1, 2) for filt in (0, 1)
         for d in ((1.0, 0.0, 0.0), (0.3, -0.5, 0.81))]


@pytest.mark.parametrize("dim,aniso,filt,direction", MODES)
def test_gabor_matches_float64(dim, aniso, filt, direction):
    n = 30
    p, dx, dy = _points(dim, n, seed=dim * 100 + aniso * 10 + filt, dscale=0.3)
    inp = np.concatenate([p, dx, dy])
    got = O.gabor(inp, anisotropic=aniso, do_filter=filt, direction=direction)[0]
    ref = np.array([_f64(p, dx, dy, i, anisotropic=aniso, do_filter=filt, direction=direction)
                    for i in range(n)])
    assert np.abs(ref).max() > 0.05  # not a trivially small field
    # measured: at most 6e-7 (values of order 0.5)
    np.testing.assert_allclose(got, ref, rtol=0, atol=3e-6)


def test_filter_is_active_where_expected():
    """Filtering changes the value once the screen footprint is wide enough
    (|dPdx x dPdy|^2 >= 1e-6) and not below it, in both implementations."""
    p, dx, dy = _points(2, 20, seed=7, dscale=0.3)
    inp = np.concatenate([p, dx, dy])
    on = O.gabor(inp, anisotropic=1, do_filter=1)[0]
    off = O.gabor(inp, anisotropic=1, do_filter=0)[0]
    assert np.mean(np.abs(on - off) > 1e-3) > 0.8
    tiny = np.concatenate([p, dx * 1e-3, dy * 1e-3])
    np.testing.assert_array_equal(O.gabor(tiny, anisotropic=1, do_filter=1)[0],
                                  O.gabor(tiny, anisotropic=1, do_filter=0)[0])


def test_layered3_configuration_and_derivatives():
    """gab = noise("gabor", Q[0], Q[1], "anisotropic", 1, "direction",
    (1,0,0), "bandwidth", 1, "impulses", 16) with Q = 8 P: value and Dual2
    derivatives."""
    rng = np.random.default_rng(21)
    n = 40
    q = rng.uniform(-8, 8, (2, n)).astype(np.float32)
    d = np.float32(8.0 * 8.0 / 1024.0)
    dx = np.stack([np.full(n, d), np.zeros(n)]).astype(np.float32)
    dy = np.stack([np.zeros(n), np.full(n, d)]).astype(np.float32)
    got = O.gabor(np.concatenate([q, dx, dy]), anisotropic=1, do_filter=1)
    ref = np.array([_f64(q, dx, dy, i, anisotropic=1, do_filter=1) for i in range(n)])
    np.testing.assert_allclose(got[0], ref, rtol=0, atol=3e-6)
    h = 1e-2
    ok = 0
    for i in range(n):
        P, DX, DY = _vec3(q, i), _vec3(dx, i), _vec3(dy, i)
        for k, dvec in ((1, DX), (2, DY)):
            fd = [(G64.gabor(P + s * dvec, DX, DY, anisotropic=1) -
                   G64.gabor(P - s * dvec, DX, DY, anisotropic=1)) / (2 * s) for s in (h, h / 2)]
            if abs(fd[0] - fd[1]) > 1e-4 + 1e-3 * abs(fd[1]):
                continue  # an impulse crossed the truncation radius
            ok += 1
            assert abs(got[k, i] - fd[1]) <= 2e-5 + 2e-3 * abs(fd[1]), (i, k, got[k, i], fd[1])
    assert ok >= 0.9 * 2 * n


def test_planted_direction_mistake_fails():
    """A transposed direction (y for x) changes the anisotropic field far
    beyond the tolerance used above."""
    p, dx, dy = _points(2, 10, seed=3, dscale=0.3)
    inp = np.concatenate([p, dx, dy])
    got = O.gabor(inp, anisotropic=1, do_filter=0, direction=(1.0, 0.0, 0.0))[0]
    wrong = np.array([_f64(p, dx, dy, i, anisotropic=1, do_filter=0, direction=(0.0, 1.0, 0.0))
                      for i in range(10)])
    assert np.abs(got - wrong).max() > 1e-2
