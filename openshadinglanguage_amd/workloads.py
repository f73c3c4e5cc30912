"""Seeded synthetic inputs of the benchmark workloads.

Every workload is a pure function of (size, seed), generated on the host
with numpy's PCG64 and handed over as float32 SoA planes [k, n].

  fbm_cell  (256x256)   P = (8u, 8v, 0) + o, u = i/255, v = j/255, one
                        seeded offset o in [0,100)^3.
  layered3  (1024x1024) P = N on the unit sphere from a stratified, jittered
                        equal-area grid; one uniform-hemisphere (about N)
                        wo/wi pair per point; dPdx = (8/1024,0,0),
                        dPdy = (0,8/1024,0).
"""
import numpy as np


def fbm_cell_inputs(res=256, seed=1):
    rng = np.random.Generator(np.random.PCG64(seed))
    o = rng.uniform(0.0, 100.0, size=3)
    idx = np.arange(res, dtype=np.float64) / (res - 1)
    u, v = np.meshgrid(idx, idx)  # v varies along rows
    P = np.stack([8.0 * u.ravel() + o[0], 8.0 * v.ravel() + o[1],
                  np.zeros(res * res) + o[2]]).astype(np.float32)
    return {"P": np.ascontiguousarray(P)}


def _hemisphere(rng, N):
    n = N.shape[1]
    a = rng.uniform(0.0, 1.0, size=n)
    b = rng.uniform(0.0, 1.0, size=n)
    z = 1.0 - 2.0 * a
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([r * np.cos(2 * np.pi * b), r * np.sin(2 * np.pi * b), z])
    flip = (d * N).sum(0) < 0
    d[:, flip] *= -1.0
    return d


def layered3_inputs(res=1024, seed=2, n=None):
    """n (optional) takes the first n points of the res x res set."""
    rng = np.random.Generator(np.random.PCG64(seed))
    j, i = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    jit = rng.uniform(0.0, 1.0, size=(2, res * res))
    u = (i.ravel() + jit[0]) / res
    v = (j.ravel() + jit[1]) / res
    z = 1.0 - 2.0 * v
    phi = 2.0 * np.pi * u
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    P = np.stack([r * np.cos(phi), r * np.sin(phi), z])
    wo = _hemisphere(rng, P)
    wi = _hemisphere(rng, P)
    sl = slice(0, n) if n is not None else slice(None)
    f = lambda a: np.ascontiguousarray(a[:, sl].astype(np.float32))  # noqa: E731
    P32 = f(P)
    return {"P": P32, "N": P32.copy(), "wo": f(wo), "wi": f(wi),
            "dPdx": np.array([8.0 / 1024, 0.0, 0.0], np.float32),
            "dPdy": np.array([0.0, 8.0 / 1024, 0.0], np.float32)}


# Layered closure material of the closure workload: white reflection tints,
# no transmission, Schlick exponent 5, base roughness = coat roughness,
# white sheen (oslmi_layered_material, include/oslmi.h).
LAYERED_MATERIAL = {
    "coat_refl_tint": (1.0, 1.0, 1.0), "coat_refr_tint": (0.0, 0.0, 0.0),
    "base_refl_tint": (1.0, 1.0, 1.0), "base_refr_tint": (0.0, 0.0, 0.0),
    "base_exponent": 5.0, "base_roughness": -1.0, "sheen_albedo": (1.0, 1.0, 1.0),
}


def layered_material():
    from ._capi import LayeredMaterial
    m = LayeredMaterial()
    for k, v in LAYERED_MATERIAL.items():
        if isinstance(v, tuple):
            getattr(m, k)[:] = list(v)
        else:
            setattr(m, k, v)
    return m


def _unit_sphere(rng, n):
    z = rng.uniform(-1.0, 1.0, size=n)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z])


def _cosine_about(rng, N):
    """Cosine-distributed directions about the unit vectors N (Malley)."""
    n = N.shape[1]
    r = np.sqrt(rng.uniform(0.0, 1.0, size=n))
    phi = rng.uniform(0.0, 2.0 * np.pi, size=n)
    x, y = r * np.cos(phi), r * np.sin(phi)
    z = np.sqrt(np.maximum(0.0, 1.0 - r * r))
    # orthonormal basis about N (Duff et al.)
    s = np.where(N[2] >= 0, 1.0, -1.0)
    a = -1.0 / (s + N[2])
    b = N[0] * N[1] * a
    T = np.stack([1.0 + s * N[0] * N[0] * a, s * b, -s * N[0]])
    B = np.stack([b, s + N[1] * N[1] * a, -N[1]])
    return x * T + y * B + z * N


def layered_closure_inputs(n=1 << 21, seed=4):
    """Closure workload: N uniform on the sphere, the view direction V
    cosine-distributed about N, the eval direction L uniform on the sphere,
    coat ior U[1.3,1.8], coat roughness U[0.05,1], sheen roughness U[0.1,1],
    F0, F90 U[0,1]^3, and 4 (rx, ry) pairs U[0,1)^2 per point."""
    rng = np.random.Generator(np.random.PCG64(seed))
    N = _unit_sphere(rng, n)
    V = _cosine_about(rng, N)
    L = _unit_sphere(rng, n)
    prm = np.concatenate([rng.uniform(1.3, 1.8, size=(1, n)), rng.uniform(0.05, 1.0, size=(1, n)),
                          rng.uniform(0.1, 1.0, size=(1, n)), rng.uniform(0.0, 1.0, size=(6, n))])
    rnd = rng.uniform(0.0, 1.0, size=(8, n))
    f = lambda a: np.ascontiguousarray(a.astype(np.float32))  # noqa: E731
    rnd32 = f(rnd)
    rnd32[rnd32 >= 1.0] = np.float32(0.99999994)
    return {"N": f(N), "V": f(V), "L": f(L), "params": f(prm), "rnd": rnd32}


def surrogate_queries(n, seed=6):
    """Closure queries in the shading frame (N = +z), 14 float32 planes:
    L xyz uniform on the sphere, V = (sqrt(1 - mu^2), 0, mu) with mu
    cosine-distributed (mu = sqrt(xi)), then the closure workload's
    parameter ranges (coat ior, coat roughness, sheen roughness, F0, F90)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = _unit_sphere(rng, n)
    mu = np.sqrt(rng.uniform(0.0, 1.0, size=n))
    V = np.stack([np.sqrt(np.maximum(0.0, 1.0 - mu * mu)), mu])
    prm = np.concatenate([rng.uniform(1.3, 1.8, size=(1, n)), rng.uniform(0.05, 1.0, size=(1, n)),
                          rng.uniform(0.1, 1.0, size=(1, n)), rng.uniform(0.0, 1.0, size=(6, n))])
    return np.ascontiguousarray(np.concatenate([L, V, prm]).astype(np.float32))
