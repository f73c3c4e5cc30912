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
