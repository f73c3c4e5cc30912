// CDNA4 device library: Gabor noise (sparse convolution with a Poisson
// impulse process, Lagae et al.), with screen-space filtering and Dual2
// derivatives.
//
// Reference: src/liboslnoise/sfm_gabornoise.h (fast_rng :21-60, kernels
// :71-97, GaborUniformParams :100-135, gabor_setup_filter :167-222,
// gabor_sample :226-300, gabor_cell :305-385, gabor_grid :390-470,
// scalar_gabor :497-515) and gabornoise.h (filter_gabor_kernel_2d
// :152-182, make_orthonormals :207-217).
//
// Per-point work that does not depend on the impulse (the filter matrices
// and, for a fixed anisotropy direction, the whole filtered-kernel
// transform) is computed once per point instead of once per impulse; the
// arithmetic on each impulse is unchanged, so values are identical.
#pragma once
#include "oslmi_math.h"
#include "oslmi_noise.h"

namespace oslmi {

// Uniform (per-op) Gabor parameters, computed on the host from NoiseParams.
struct GaborUniform {
    float a, lambda, sqrt_lambda_inv, radius, radius2, radius3, radius_inv;
    float poisson_g;  // expf(-lambda*radius3)
    float scale;      // final normalization of scalar_gabor
    float dir_x, dir_y, dir_z;
    int anisotropic;  // 0 isotropic, 1 anisotropic, 2 hybrid
    int do_filter;
};

struct M22 {
    float m00, m01, m10, m11;
};
OSLMI_DEV M22 m22(float a, float b, float c, float d)
{
    M22 r;
    r.m00 = a;
    r.m01 = b;
    r.m10 = c;
    r.m11 = d;
    return r;
}
OSLMI_DEV M22 mul(const M22& A, const M22& B)
{
    return m22(0.0f + A.m00 * B.m00 + A.m01 * B.m10, 0.0f + A.m00 * B.m01 + A.m01 * B.m11,
               0.0f + A.m10 * B.m00 + A.m11 * B.m10, 0.0f + A.m10 * B.m01 + A.m11 * B.m11);
}
OSLMI_DEV M22 add(const M22& A, const M22& B)
{
    return m22(A.m00 + B.m00, A.m01 + B.m01, A.m10 + B.m10, A.m11 + B.m11);
}
OSLMI_DEV M22 scale(float s, const M22& A) { return m22(A.m00 * s, A.m01 * s, A.m10 * s, A.m11 * s); }
OSLMI_DEV float det(const M22& A) { return A.m00 * A.m11 - A.m01 * A.m10; }
// Imath Matrix22::inverse(singExc=false)
OSLMI_DEV M22 inv(const M22& A)
{
    float s[4] = { A.m11, -A.m01, -A.m10, A.m00 };
    float r    = A.m00 * A.m11 - A.m10 * A.m01;
    if (fabsf(r) >= 1.0f) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            s[i] /= r;
    } else {
        float mr = fabsf(r) / 1.17549435e-38f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (mr > fabsf(s[i]))
                s[i] /= r;
            else
                return m22(1.0f, 0.0f, 0.0f, 1.0f);
        }
    }
    return m22(s[0], s[1], s[2], s[3]);
}

struct M33 {
    float m[3][3];
};

OSLMI_DEV V3 mult(const M33& M, V3 s)
{
    return V3(s.x * M.m[0][0] + s.y * M.m[1][0] + s.z * M.m[2][0],
              s.x * M.m[0][1] + s.y * M.m[1][1] + s.z * M.m[2][1],
              s.x * M.m[0][2] + s.y * M.m[1][2] + s.z * M.m[2][2]);
}
// Columns 0 and 1 of (row vector s) * M, with derivatives.
OSLMI_DEV void mult_xy(const M33& M, const Dv& s, Df& a, Df& b)
{
    Df s0 = comp(s, 0), s1 = comp(s, 1), s2 = comp(s, 2);
    a     = s0 * M.m[0][0] + s1 * M.m[1][0] + s2 * M.m[2][0];
    b     = s0 * M.m[0][1] + s1 * M.m[1][1] + s2 * M.m[2][1];
}

OSLMI_DEV V3 vnormalized(V3 v)
{
    float l = sqrtf(dot(v, v));
    return l != 0.0f ? V3(v.x / l, v.y / l, v.z / l) : v;
}

OSLMI_DEV Df ddot(const Dv& a, const Dv& b)
{
    return comp(a, 0) * comp(b, 0) + comp(a, 1) * comp(b, 1) + comp(a, 2) * comp(b, 2);
}
OSLMI_DEV Df ddot(V3 a, const Dv& b) { return comp(b, 0) * a.x + comp(b, 1) * a.y + comp(b, 2) * a.z; }

struct GaborRng {
    uint32_t s;
    OSLMI_DEV GaborRng(V3 p, uint32_t seed)
    {
        s = inthash((uint32_t)ifloor(p.x), (uint32_t)ifloor(p.y), (uint32_t)ifloor(p.z), seed);
        if (!s)
            s = 1;
    }
    OSLMI_DEV float operator()()
    {
        s *= 3039177861u;
        return (float)s * (1.0f / 4294967296.0f);
    }
};

// Per-point filter state (gabor_setup_filter) plus, for the anisotropic
// mode, the impulse-independent part of filter_gabor_kernel_2d.
struct GaborPoint {
    M22 filter;
    M33 local;
    V3 N;
    int do_filter;
    // anisotropic + filter: filtered-kernel constants
    V3 omega_t;
    float c_F, K, E, a_f, omf_x, omf_y;
};

OSLMI_DEV void gabor_filter_consts(const M22& Sigma_f, float a, float ox, float oy, float& c_F,
                                   float& K, float& E, float& a_f, float& omf_x, float& omf_y)
{
    const float two_pi = 6.28318548202514648438f;  // float(2*pi)
    M22 Sigma_G        = scale(a * a / two_pi, m22(1.0f, 0.0f, 0.0f, 1.0f));
    c_F                = 1.0f / (two_pi * sqrtf(det(Sigma_f)));
    M22 Sigma_F        = scale((float)(1.0 / (4.0 * 3.14159265358979323846 * 3.14159265358979323846)),
                               inv(Sigma_f));
    M22 SGF            = add(Sigma_G, Sigma_F);
    M22 SGFi           = inv(SGF);
    float mx           = ox * SGFi.m00 + oy * SGFi.m10;
    float my           = ox * SGFi.m01 + oy * SGFi.m11;
    E                  = fexp(-0.5f * (mx * ox + my * oy));
    K                  = 1.0f / (two_pi * sqrtf(det(SGF)));
    M22 Sigma_G_i      = inv(Sigma_G);
    M22 Sigma_GF       = inv(add(inv(Sigma_F), Sigma_G_i));
    M22 Sigma_GF_Gi    = mul(Sigma_GF, Sigma_G_i);
    omf_x              = ox * Sigma_GF_Gi.m00 + oy * Sigma_GF_Gi.m10;
    omf_y              = ox * Sigma_GF_Gi.m01 + oy * Sigma_GF_Gi.m11;
    a_f = sqrtf((float)(6.283185307179586476925 * (double)sqrtf(det(Sigma_GF))));
}

OSLMI_DEV void gabor_setup(const GaborUniform& U, const Dv& P, GaborPoint& gp)
{
    gp.do_filter = 0;
    if (!U.do_filter)
        return;
    V3 n = cross(P.dx, P.dy);
    if (dot(n, n) < 1.0e-6f)
        return;
    n = vnormalized(n);
    V3 t = fabsf(n.x) < 0.9f ? V3(0.0f, n.z, -n.y) : V3(-n.z, 0.0f, n.x);
    t    = vnormalized(t);
    V3 b = cross(n, t);
    // Mtex_to_tan = cols(t, b, n); Mscreen_to_tex = cols(dPdx, dPdy, 0)
    M33 L;
    L.m[0][0] = t.x; L.m[0][1] = b.x; L.m[0][2] = n.x;
    L.m[1][0] = t.y; L.m[1][1] = b.y; L.m[1][2] = n.y;
    L.m[2][0] = t.z; L.m[2][1] = b.z; L.m[2][2] = n.z;
    // rows 0,1 / cols 0,1 of Mscreen_to_tex * Mtex_to_tan
    float s00 = 0.0f + P.dx.x * L.m[0][0] + P.dy.x * L.m[1][0] + 0.0f * L.m[2][0];
    float s01 = 0.0f + P.dx.x * L.m[0][1] + P.dy.x * L.m[1][1] + 0.0f * L.m[2][1];
    float s10 = 0.0f + P.dx.y * L.m[0][0] + P.dy.y * L.m[1][0] + 0.0f * L.m[2][0];
    float s11 = 0.0f + P.dx.y * L.m[0][1] + P.dy.y * L.m[1][1] + 0.0f * L.m[2][1];
    M22 Mst   = m22(s00, s01, s10, s11);
    M22 Mt    = m22(s00, s10, s01, s11);
    float sf  = 0.5f;
    M22 Sf    = mul(mul(Mt, m22(sf * sf, 0.0f, 0.0f, sf * sf)), Mst);
    gp.N      = n;
    gp.filter = Sf;
    gp.local  = L;
    if (det(Sf) < 1.0e-18f)
        return;
    gp.do_filter = 1;
    if (U.anisotropic == 1) {
        gp.omega_t = mult(L, V3(U.dir_x, U.dir_y, U.dir_z));
        gabor_filter_consts(Sf, U.a, gp.omega_t.x, gp.omega_t.y, gp.c_F, gp.K, gp.E, gp.a_f,
                            gp.omf_x, gp.omf_y);
    }
}

OSLMI_DEV Df gabor_kernel3(V3 omega, float phi, float bw, const Dv& x)
{
    const float pi = 3.14159274101257324219f, two_pi = 6.28318548202514648438f;
    Df g = dexp((-pi * (bw * bw)) * ddot(x, x));
    Df h = dcos(two_pi * ddot(omega, x) + Df(phi));
    return Df(1.0f) * g * h;
}
OSLMI_DEV Df gabor_kernel2(const Df& w, float ox, float oy, const Df& phi, float bw, const Df& x0,
                           const Df& x1)
{
    const float pi = 3.14159274101257324219f, two_pi = 6.28318548202514648438f;
    Df g = dexp((-pi * (bw * bw)) * (x0 * x0 + x1 * x1));
    Df h = dcos(two_pi * (x0 * ox + x1 * oy) + phi);
    return w * g * h;
}

// One in-radius impulse: contribution to the cell sum.
OSLMI_DEV Df gabor_impulse(const GaborUniform& U, const GaborPoint& gp, V3 omega, float phi,
                           const Dv& xk)
{
    const float pi = 3.14159274101257324219f, two_pi = 6.28318548202514648438f;
    if (gp.do_filter == 1) {
        V3 omega_t;
        float c_F, K, E, a_f, ofx, ofy;
        if (U.anisotropic == 1) {
            omega_t = gp.omega_t;
            c_F = gp.c_F; K = gp.K; E = gp.E; a_f = gp.a_f; ofx = gp.omf_x; ofy = gp.omf_y;
        } else {
            omega_t = mult(gp.local, omega);
            gabor_filter_consts(gp.filter, U.a, omega_t.x, omega_t.y, c_F, K, E, a_f, ofx, ofy);
        }
        Df d_i   = -ddot(gp.N, xk);
        Df w_s   = 1.0f * dexp((-pi * (U.a * U.a)) * (d_i * d_i));
        Df phi_s = phi - two_pi * d_i * omega_t.z;
        Df w_f   = (c_F * w_s) * K * E;
        Df x0, x1;
        mult_xy(gp.local, xk, x0, x1);
        Df gk = gabor_kernel2(w_f, ofx, ofy, phi_s, a_f, x0, x1);
        if (!isfinite(gk.v))
            gk = gabor_kernel3(omega, phi, U.a, xk);
        return gk;
    }
    return gabor_kernel3(omega, phi, U.a, xk);
}

// gabor_cell: impulses of one lattice cell, summed in generation order.
// seed selects the generator stream (0 for the scalar noise, 0/1/2 for the
// components of the vector noise); a periodic noise wraps the cell corner
// before it seeds the generator (sfm_gabornoise.h:304-310).
OSLMI_DEV Df gabor_cell(const GaborUniform& U, const GaborPoint& gp, V3 c_i, const Dv& x_c_i,
                        uint32_t seed, bool periodic, V3 period)
{
    const float two_pi = 6.28318548202514648438f;
    if (periodic)
        c_i = V3(wrap_period(c_i.x, period.x), wrap_period(c_i.y, period.y),
                 wrap_period(c_i.z, period.z));
    GaborRng rng(c_i, seed);
    // Poisson(lambda*radius3) by multiplying uniforms (fast_rng::poisson)
    int n = 0;
    {
        float t = rng();
        while (t > U.poisson_g) {
            ++n;
            t *= rng();
        }
    }
    Df sum = Df(0.0f);
    for (int i = 0; i < n; ++i) {
        float z_rng = rng(), y_rng = rng(), x_rng = rng();
        V3 x_i_c(x_rng, y_rng, z_rng);
        Dv xk = U.radius * Dv(x_c_i.v - x_i_c, x_c_i.dx, x_c_i.dy);
        if (dot(xk.v, xk.v) < U.radius2) {
            V3 omega;
            if (U.anisotropic == 1) {
                omega = V3(U.dir_x, U.dir_y, U.dir_z);
            } else if (U.anisotropic == 0) {
                float omega_t = two_pi * rng();
                float u       = rng();
                float cos_p   = -1.0f * (1.0f - u) + 1.0f * u;
                float sin_p   = sqrtf(fmaxf(0.0f, 1.0f - cos_p * cos_p));
                float st, ct;
                fsincos(omega_t, &st, &ct);
                omega = vnormalized(V3(ct * sin_p, st * sin_p, cos_p));
            } else {
                float omega_r = sqrtf(U.dir_x * U.dir_x + U.dir_y * U.dir_y + U.dir_z * U.dir_z);
                float omega_t = two_pi * rng();
                float st, ct;
                fsincos(omega_t, &st, &ct);
                omega = omega_r * V3(ct, st, 0.0f);
            }
            float phi = two_pi * rng();
            sum       = sum + gabor_impulse(U, gp, omega, phi, xk);
        } else {
            if (U.anisotropic == 0) {
                rng();
                rng();
            } else if (U.anisotropic != 1) {
                rng();
            }
            rng();
        }
    }
    return sum;
}

// gabor_evaluate + gabor_grid: the 27 cells around P, scaled.
OSLMI_DEV Df gabor_grid(const GaborUniform& U, const GaborPoint& gp, const Dv& P, uint32_t seed,
                        bool periodic, V3 period)
{
    Dv xg = P * U.radius_inv;
    V3 fl(floorf(xg.v.x), floorf(xg.v.y), floorf(xg.v.z));
    Dv xc(xg.v - fl, xg.dx, xg.dy);
    Df sum = Df(0.0f);
    for (int k = -1; k <= 1; ++k)
        for (int j = -1; j <= 1; ++j)
            for (int i = -1; i <= 1; ++i) {
                V3 c((float)i, (float)j, (float)k);
                sum = sum + gabor_cell(U, gp, fl + c, Dv(xc.v - c, xc.dx, xc.dy), seed, periodic,
                                       period);
            }
    return (sum * U.sqrt_lambda_inv) * U.scale;
}

// scalar_gabor / scalar_pgabor for a Dual2 point P (dim 1..3 input already
// packed in P; pgabor pads the periods of 1 and 2 dimensions with zeros).
OSLMI_DEV Df gabor(const GaborUniform& U, const Dv& P, bool periodic = false,
                   V3 period = V3(0.0f))
{
    GaborPoint gp;
    gabor_setup(U, P, gp);
    return gabor_grid(U, gp, P, 0u, periodic, period);
}

// scalar_gabor3 / scalar_pgabor3 (sfm_gabornoise.h:517-640): one filter
// set-up, the three components from generator seeds 0, 1 and 2. The seed
// loop is kept rolled: the cell walk is large and runs once per component.
OSLMI_DEV Dv gabor3(const GaborUniform& U, const Dv& P, bool periodic = false,
                    V3 period = V3(0.0f))
{
    GaborPoint gp;
    gabor_setup(U, P, gp);
    Df r0, r1, r2;
#pragma clang loop unroll(disable)
    for (uint32_t seed = 0; seed < 3; ++seed) {
        const Df r = gabor_grid(U, gp, P, seed, periodic, period);
        if (seed == 0)
            r0 = r;
        else if (seed == 1)
            r1 = r;
        else
            r2 = r;
    }
    return make_dv(r0, r1, r2);
}

}  // namespace oslmi
