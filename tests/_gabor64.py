"""Gabor noise restated in float64 numpy, as a check of the oracle's binary32
implementation (Lagae et al. 2009/2011 sparse convolution noise, as
src/liboslnoise/sfm_gabornoise.h and gabornoise.h evaluate it).

The discrete decisions are made exactly as the reference makes them, in
binary32: the cell of the point, the fast_rng stream seeded by the cell's
lookup3 hash (fast_rng, sfm_gabornoise.h:21-60: multiplier 3039177861,
uniforms u32 * 2^-32), the Poisson impulse count against expf(-mean), the
impulse order (z, y, x draws) and the in-radius test. Everything that is
then summed (kernel envelope and harmonic, the 2D slice of section 3.2 and
the Gaussian filter convolution of Equation 10, the output scale) is
evaluated in float64 from the same formulas, with numpy's linear algebra
in place of the reference's hand-written 2x2 algebra. A wrong direction,
frequency, phase, slice or filter term therefore changes the result by
orders of magnitude more than binary32 rounding does.
"""
import numpy as np

import _oracle as O

F32 = np.float32
TWO_PI = 2.0 * np.pi


def uniform_params(bandwidth=1.0, impulses=16.0):
    """GaborUniformParams (sfm_gabornoise.h:100-135) in binary32 (the
    Poisson mean and the radius test use these) and float64."""
    bw = min(max(bandwidth, 0.01), 100.0)
    two_bw = 2.0 ** bw
    a = 2.0 * ((two_bw - 1.0) / (two_bw + 1.0)) * 2.128934
    radius = np.sqrt(-np.log(0.02) / np.pi) / a
    imp = min(max(impulses, 1.0), 32.0)
    lam = imp / (1.33333 * np.pi * radius ** 3)
    # binary32 copies in the reference's operation order
    a32 = F32(F32(2.0) * (F32(F32(two_bw) - F32(1)) / F32(F32(two_bw) + F32(1)))
              * F32(2.128934))
    r32 = F32(np.sqrt(F32(-np.log(F32(0.02)) / F32(np.pi)), dtype=F32) / a32)
    r2_32 = F32(r32 * r32)
    r3_32 = F32(r2_32 * r32)
    lam32 = F32(F32(imp) / F32(F32(1.33333 * np.pi) * r3_32))
    return {"a": a, "radius": radius, "lambda": lam, "a32": a32, "r32": r32,
            "r2_32": r2_32, "rinv32": F32(F32(1) / r32), "mean32": F32(lam32 * r3_32)}


class Rng:
    def __init__(self, cell, seed):
        s = O.inthash([int(np.int64(c)) & 0xffffffff for c in cell] + [seed & 0xffffffff])
        self.s = s if s else 1

    def __call__(self):
        self.s = (self.s * 3039177861) & 0xffffffff
        return F32(F32(self.s) * F32(2.0 ** -32))

    def poisson(self, mean32):
        g = np.exp(-mean32, dtype=F32)
        em, t = 0, self()
        while t > g:
            em += 1
            t = F32(t * self())
        return em


def orthonormals(v):
    v = v / np.linalg.norm(v)
    a = np.array([0.0, v[2], -v[1]]) if abs(v[0]) < 0.9 else np.array([-v[2], 0.0, v[0]])
    a = a / np.linalg.norm(a)
    return v, a, np.cross(v, a)


def setup_filter(dPdx, dPdy):
    """gabor_setup_filter (sfm_gabornoise.h:167-222): the screen-space
    filter (sigma 1/2) carried into the tangent frame of the slice."""
    n = np.cross(dPdx, dPdy)
    if n.dot(n) < 1e-6:
        return None
    n, t, b = orthonormals(n)
    local = np.stack([t, b, n], axis=1)            # columns t, b, n
    scr_tex = np.stack([dPdx, dPdy, np.zeros(3)], axis=1)
    m = (scr_tex @ local)[:2, :2]
    sigma = m.T @ (0.25 * np.eye(2)) @ m
    if np.linalg.det(sigma) < 1e-18:
        return None
    return {"N": n, "local": local, "filter": sigma}


def filtered_kernel(filt, w, a, omega, phi):
    """filter_gabor_kernel_2d (gabornoise.h:152-182): the Gaussian-envelope
    harmonic convolved with the Gaussian filter, again a Gabor kernel."""
    sig_f = filt
    sig_g = (a * a / TWO_PI) * np.eye(2)
    c_f = 1.0 / (TWO_PI * np.sqrt(np.linalg.det(sig_f)))
    sig_F = np.linalg.inv(sig_f) / (4.0 * np.pi * np.pi)
    s = sig_g + sig_F
    c_gf = c_f * w / (TWO_PI * np.sqrt(np.linalg.det(s))) * np.exp(
        -0.5 * (omega @ np.linalg.inv(s)) @ omega)
    sig_gf = np.linalg.inv(np.linalg.inv(sig_F) + np.linalg.inv(sig_g))
    mu = omega @ (sig_gf @ np.linalg.inv(sig_g))
    a_f = np.sqrt(TWO_PI * np.sqrt(np.linalg.det(sig_gf)))
    return c_gf, a_f, mu


def gabor(P, dPdx, dPdy, anisotropic=0, do_filter=1, direction=(1.0, 0.0, 0.0),
          bandwidth=1.0, impulses=16.0, seed=0):
    """scalar_gabor (sfm_gabornoise.h:497-515) at one point: P, dPdx, dPdy
    float64 3-vectors (the binary32 values of the shading point)."""
    up = uniform_params(bandwidth, impulses)
    a, radius = up["a"], up["radius"]
    direction = np.asarray(direction, np.float64)
    filt = setup_filter(dPdx, dPdy) if do_filter else None
    # binary32 cell search (the reference's x_g = x * radius_inv)
    xg32 = (np.asarray(P, F32) * up["rinv32"]).astype(F32)
    fl32 = np.floor(xg32)
    xc32 = (xg32 - fl32).astype(F32)
    xg = np.asarray(P, np.float64) / radius
    xc = xg - fl32.astype(np.float64)
    total = 0.0
    for k in (-1, 0, 1):
        for j in (-1, 0, 1):
            for i in (-1, 0, 1):
                c = np.array([i, j, k], F32)
                cell = (fl32 + c).astype(F32)
                rng = Rng([int(np.floor(v)) for v in cell], seed)
                xci32 = (xc32 - c).astype(F32)
                xci = xc - c.astype(np.float64)
                for _ in range(rng.poisson(up["mean32"])):
                    z, y, x = rng(), rng(), rng()
                    imp = np.array([x, y, z], F32)
                    xk32 = (up["r32"] * (xci32 - imp)).astype(F32)
                    if not F32(F32(F32(xk32[0] * xk32[0]) + F32(xk32[1] * xk32[1]))
                               + F32(xk32[2] * xk32[2])) < up["r2_32"]:
                        rng()  # gabor_no_sample: the phase draw, and the
                        if anisotropic == 0:  # orientation draws
                            rng(), rng()
                        elif anisotropic == 2:
                            rng()
                        continue
                    xk = radius * (xci - imp.astype(np.float64))
                    if anisotropic == 1:
                        omega = direction
                    elif anisotropic == 0:
                        ot = TWO_PI * float(rng())
                        cp = -1.0 + 2.0 * float(rng())
                        sp = np.sqrt(max(0.0, 1.0 - cp * cp))
                        omega = np.array([np.cos(ot) * sp, np.sin(ot) * sp, cp])
                        omega = omega / np.linalg.norm(omega)
                    else:
                        ot = TWO_PI * float(rng())
                        omega = np.linalg.norm(direction) * np.array([np.cos(ot), np.sin(ot), 0.0])
                    phi = TWO_PI * float(rng())
                    if filt is not None:
                        ot3 = omega @ filt["local"]
                        d = -filt["N"].dot(xk)
                        w_s = np.exp(-np.pi * a * a * d * d)
                        phi_s = phi - TWO_PI * d * ot3[2]
                        w_f, a_f, om_f = filtered_kernel(filt["filter"], w_s, a, ot3[:2], phi_s)
                        xt = (xk @ filt["local"])[:2]
                        total += w_f * np.exp(-np.pi * a_f * a_f * xt.dot(xt)) * np.cos(
                            TWO_PI * om_f.dot(xt) + phi_s)
                    else:
                        total += np.exp(-np.pi * a * a * xk.dot(xk)) * np.cos(
                            TWO_PI * omega.dot(xk) + phi)
    variance = 1.0 / (4.0 * np.sqrt(2.0) * a ** 3)
    scale = 0.5 / (3.0 * np.sqrt(variance))
    return total / np.sqrt(up["lambda"]) * scale
