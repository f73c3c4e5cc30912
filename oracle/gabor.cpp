// CPU oracle: Gabor noise.
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
//
// Follows src/liboslnoise/sfm_gabornoise.h (the batched path's "SIMD
// friendly" scalar variant) and gabornoise.h:
//   fast_rng                 sfm_gabornoise.h:21-60
//   gabor_kernel             sfm_gabornoise.h:71-80
//   slice_gabor_kernel_3d    sfm_gabornoise.h:82-97
//   GaborUniformParams       sfm_gabornoise.h:100-135
//   gabor_setup_filter       sfm_gabornoise.h:167-222
//   gabor_sample / no_sample sfm_gabornoise.h:226-300
//   gabor_cell               sfm_gabornoise.h:305-385
//   gabor_grid               sfm_gabornoise.h:390-470
//   scalar_gabor             sfm_gabornoise.h:497-515
//   filter_gabor_kernel_2d   gabornoise.h:152-182
//   make_orthonormals        gabornoise.h:207-217
// Imath Matrix22/Matrix33 arithmetic (determinant, inverse, products) is
// restated from the Imath 3 definitions; Imath's source is not part of the
// reference tree. exp and sincos use the oracle's fixed approximations.
#include "gabor.h"

#include <algorithm>
#include <cfloat>

#include "noise.h"
#include "omath.h"

namespace oref {

namespace {

const float kPi    = float(M_PI);
const float kTwoPi = float(2.0 * M_PI);

struct M22 {
    float m[2][2];
    M22() { m[0][0] = 1; m[0][1] = 0; m[1][0] = 0; m[1][1] = 1; }
    M22(float a, float b, float c, float d) { m[0][0] = a; m[0][1] = b; m[1][0] = c; m[1][1] = d; }
};
M22 operator*(const M22& A, const M22& B)
{
    M22 r(0, 0, 0, 0);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            float acc = 0.0f;
            for (int k = 0; k < 2; ++k)
                acc += A.m[i][k] * B.m[k][j];
            r.m[i][j] = acc;
        }
    return r;
}
M22 operator+(const M22& A, const M22& B)
{
    return M22(A.m[0][0] + B.m[0][0], A.m[0][1] + B.m[0][1], A.m[1][0] + B.m[1][0],
               A.m[1][1] + B.m[1][1]);
}
M22 operator*(float s, const M22& A)
{
    return M22(A.m[0][0] * s, A.m[0][1] * s, A.m[1][0] * s, A.m[1][1] * s);
}
M22 transposed(const M22& A) { return M22(A.m[0][0], A.m[1][0], A.m[0][1], A.m[1][1]); }
// Imathx.h:89-94
float determinant(const M22& A) { return A.m[0][0] * A.m[1][1] - A.m[0][1] * A.m[1][0]; }
// Imath Matrix22::inverse(singExc=false): adjugate divided elementwise by
// the determinant, with the underflow guard for |det| < 1.
M22 inverse(const M22& A)
{
    M22 s(A.m[1][1], -A.m[0][1], -A.m[1][0], A.m[0][0]);
    float r = A.m[0][0] * A.m[1][1] - A.m[1][0] * A.m[0][1];
    if (fabsf(r) >= 1.0f) {
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                s.m[i][j] /= r;
    } else {
        float mr = fabsf(r) / FLT_MIN;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                if (mr > fabsf(s.m[i][j]))
                    s.m[i][j] /= r;
                else
                    return M22();
            }
    }
    return s;
}

struct M33 {
    float m[3][3];
};
M33 cols(V3 a, V3 b, V3 c)
{
    M33 r;
    r.m[0][0] = a.x; r.m[0][1] = b.x; r.m[0][2] = c.x;
    r.m[1][0] = a.y; r.m[1][1] = b.y; r.m[1][2] = c.y;
    r.m[2][0] = a.z; r.m[2][1] = b.z; r.m[2][2] = c.z;
    return r;
}
M33 operator*(const M33& A, const M33& B)
{
    M33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float acc = 0.0f;
            for (int k = 0; k < 3; ++k)
                acc += A.m[i][k] * B.m[k][j];
            r.m[i][j] = acc;
        }
    return r;
}
// multMatrix (Imathx.h:32-42): row vector times matrix
V3 mult(const M33& M, V3 s)
{
    return V3(s.x * M.m[0][0] + s.y * M.m[1][0] + s.z * M.m[2][0],
              s.x * M.m[0][1] + s.y * M.m[1][1] + s.z * M.m[2][1],
              s.x * M.m[0][2] + s.y * M.m[1][2] + s.z * M.m[2][2]);
}
// multMatrix with derivs (dual_vec.h:288-299)
D2<V3> mult(const M33& M, const D2<V3>& s)
{
    D2<float> s0(s.v.x, s.dx.x, s.dy.x), s1(s.v.y, s.dx.y, s.dy.y), s2(s.v.z, s.dx.z, s.dy.z);
    D2<float> a = s0 * M.m[0][0] + s1 * M.m[1][0] + s2 * M.m[2][0];
    D2<float> b = s0 * M.m[0][1] + s1 * M.m[1][1] + s2 * M.m[2][1];
    D2<float> c = s0 * M.m[0][2] + s1 * M.m[1][2] + s2 * M.m[2][2];
    return D2<V3>(V3(a.v, b.v, c.v), V3(a.dx, b.dx, c.dx), V3(a.dy, b.dy, c.dy));
}

// Imath Vec3::length / normalize
float length(V3 v) { return sqrtf(dot(v, v)); }
V3 normalized(V3 v)
{
    float l = length(v);
    if (l != 0.0f)
        return V3(v.x / l, v.y / l, v.z / l);
    return v;
}

D2<float> dexp(const D2<float>& a)
{
    float f = ofast_expf(a.v);
    return D2<float>(f, f * a.dx, f * a.dy);
}
D2<float> dcos(const D2<float>& a)
{
    float s, c;
    ofast_sincosf(a.v, &s, &c);
    return D2<float>(c, -s * a.dx, -s * a.dy);
}
D2<float> comp(const D2<V3>& a, int i)
{
    return i == 0 ? D2<float>(a.v.x, a.dx.x, a.dy.x)
                  : (i == 1 ? D2<float>(a.v.y, a.dx.y, a.dy.y) : D2<float>(a.v.z, a.dx.z, a.dy.z));
}
// dot(Dual<Vec3>, Dual<Vec3>) and dot(Vec3, Dual<Vec3>) (dual_vec.h:412-445)
D2<float> ddot(const D2<V3>& a, const D2<V3>& b)
{
    return comp(a, 0) * comp(b, 0) + comp(a, 1) * comp(b, 1) + comp(a, 2) * comp(b, 2);
}
D2<float> ddot(V3 a, const D2<V3>& b)
{
    return comp(b, 0) * a.x + comp(b, 1) * a.y + comp(b, 2) * a.z;
}

struct FastRng {
    uint32_t s;
    FastRng(V3 p, int seed)
    {
        s = inthash((uint32_t)ifloor(p.x), (uint32_t)ifloor(p.y), (uint32_t)ifloor(p.z),
                    (uint32_t)seed);
        if (!s)
            s = 1;
    }
    float operator()()
    {
        s *= 3039177861u;
        return (float)s * (1.0f / 4294967296.0f);
    }
    int poisson(float g)
    {
        unsigned em = 0;
        float t     = (*this)();
        while (t > g) {
            ++em;
            t *= (*this)();
        }
        return (int)em;
    }
};

// filter_gabor_kernel_2d split in two: the part that depends only on the
// filter, the bandwidth and the impulse's 2D frequency (FilterConsts), and
// the per-impulse weight  w_f = ((c_F * w) * K) * E  (filter_weight). With a
// fixed anisotropy direction the first part is the same for every impulse
// of a point and is evaluated once per point; the values are unchanged.
struct FilterConsts {
    float c_F, K, E, a_f, omf_x, omf_y;
};

FilterConsts filter_consts(const M22& filter, float a, float ox, float oy)
{
    FilterConsts f;
    const M22& Sigma_f = filter;
    M22 Sigma_G        = (a * a / kTwoPi) * M22();
    f.c_F              = 1.0f / (kTwoPi * sqrtf(determinant(Sigma_f)));
    M22 Sigma_F        = float(1.0 / (4.0 * M_PI * M_PI)) * inverse(Sigma_f);
    M22 SGF            = Sigma_G + Sigma_F;
    M22 SGFi           = inverse(SGF);
    // gabor_mul_m22_v2(M, v) = (v.x*M00 + v.y*M10, v.x*M01 + v.y*M11)
    float mx = ox * SGFi.m[0][0] + oy * SGFi.m[1][0];
    float my = ox * SGFi.m[0][1] + oy * SGFi.m[1][1];
    f.E      = ofast_expf(-0.5f * (mx * ox + my * oy));
    f.K      = 1.0f / (kTwoPi * sqrtf(determinant(SGF)));
    M22 Sigma_G_i   = inverse(Sigma_G);
    M22 Sigma_GF    = inverse(inverse(Sigma_F) + Sigma_G_i);
    M22 Sigma_GF_Gi = Sigma_GF * Sigma_G_i;
    f.omf_x         = ox * Sigma_GF_Gi.m[0][0] + oy * Sigma_GF_Gi.m[1][0];
    f.omf_y         = ox * Sigma_GF_Gi.m[0][1] + oy * Sigma_GF_Gi.m[1][1];
    f.a_f           = sqrtf((float)(2.0 * M_PI * (double)sqrtf(determinant(Sigma_GF))));
    return f;
}

D2<float> filter_weight(const FilterConsts& f, const D2<float>& w)
{
    return (f.c_F * w) * f.K * f.E;
}

struct Params {
    M22 filter;
    M33 local;
    V3 N;
    V3 omega;
    int do_filter;
    V3 omega_t;       // anisotropic: local * omega
    FilterConsts fc;  // anisotropic: filter_consts(filter, a, omega_t.xy)
};

// gabor_kernel with a 3D (dual) position
D2<float> kernel3(const D2<float>& w, V3 omega, const D2<float>& phi, float bw, const D2<V3>& x)
{
    D2<float> g = dexp((float(-kPi) * (bw * bw)) * ddot(x, x));
    D2<float> h = dcos(kTwoPi * ddot(omega, x) + phi);
    return w * g * h;
}
// gabor_kernel with a 2D (dual) position given as two components
D2<float> kernel2(const D2<float>& w, float ox, float oy, const D2<float>& phi, float bw,
                  const D2<float>& x0, const D2<float>& x1)
{
    D2<float> g = dexp((float(-kPi) * (bw * bw)) * (x0 * x0 + x1 * x1));
    D2<float> h = dcos(kTwoPi * (x0 * ox + x1 * oy) + phi);
    return w * g * h;
}

void setup_filter(const D2<V3>& P, Params& gp)
{
    V3 n          = cross(P.dx, P.dy);
    int do_filter = 1;
    if (dot(n, n) < 1.0e-6f) {
        do_filter = 0;
    } else {
        n = normalized(n);  // make_orthonormals
        V3 t;
        if (fabsf(n.x) < 0.9f)
            t = V3(0.0f, n.z, -n.y);
        else
            t = V3(-n.z, 0.0f, n.x);
        t          = normalized(t);
        V3 b       = cross(n, t);
        M33 tex2tan = cols(t, b, n);
        M33 scr2tex = cols(P.dx, P.dy, V3(0.0f, 0.0f, 0.0f));
        M33 scr2tan = scr2tex * tex2tan;
        M22 Mst(scr2tan.m[0][0], scr2tan.m[0][1], scr2tan.m[1][0], scr2tan.m[1][1]);
        float sf = 0.5f;
        M22 Sigma_f_scr(sf * sf, 0.0f, 0.0f, sf * sf);
        M22 Sigma_f_tan = transposed(Mst) * Sigma_f_scr * Mst;
        gp.N            = n;
        gp.filter       = Sigma_f_tan;
        gp.local        = tex2tan;
        if (determinant(Sigma_f_tan) < 1.0e-18f)
            do_filter = 0;
    }
    gp.do_filter = do_filter;
}

// gabor_cell (sfm_gabornoise.h:304-380): with a period the cell's corner
// is wrapped (gabornoise.h:185-199) before it seeds the generator.
D2<float> cell(const GaborUniform& gup, const GaborOpts& opt, const Params& gp, V3 c_i,
               const D2<V3>& x_c_i, int seed, const V3* period)
{
    if (period)
        c_i = V3(wrap_period(c_i.x, period->x), wrap_period(c_i.y, period->y),
                 wrap_period(c_i.z, period->z));
    FastRng rng(c_i, seed);
    int n_imp     = rng.poisson(gup.poisson_g);
    D2<float> sum = D2<float>(0.0f);
    for (int i = 0; i < n_imp; ++i) {
        float z_rng = rng(), y_rng = rng(), x_rng = rng();
        V3 x_i_c(x_rng, y_rng, z_rng);
        D2<V3> d  = D2<V3>(x_c_i.v - x_i_c, x_c_i.dx, x_c_i.dy);
        D2<V3> xk = gup.radius * d;
        if (dot(xk.v, xk.v) < gup.radius2) {
            V3 omega;
            float phi;
            if (opt.anisotropic == 1) {
                omega = gp.omega;
            } else if (opt.anisotropic == 0) {
                float omega_t = kTwoPi * rng();
                float u       = rng();
                float cos_p   = -1.0f * (1.0f - u) + 1.0f * u;  // OIIO::lerp(-1,1,u)
                float sin_p   = sqrtf(std::max<float>(0.0f, 1.0f - cos_p * cos_p));
                float st, ct;
                ofast_sincosf(omega_t, &st, &ct);
                omega = normalized(V3(ct * sin_p, st * sin_p, cos_p));
            } else {
                float omega_r = length(gp.omega);
                float omega_t = kTwoPi * rng();
                float st, ct;
                ofast_sincosf(omega_t, &st, &ct);
                omega = omega_r * V3(ct, st, 0.0f);
            }
            phi = kTwoPi * rng();
            if (opt.do_filter && gp.do_filter == 1) {
                V3 omega_t;
                FilterConsts fc;
                if (opt.anisotropic == 1) {
                    omega_t = gp.omega_t;
                    fc      = gp.fc;
                } else {
                    omega_t = mult(gp.local, omega);
                    fc      = filter_consts(gp.filter, gup.a, omega_t.x, omega_t.y);
                }
                D2<float> d_i = -ddot(gp.N, xk);
                // slice_gabor_kernel_3d
                D2<float> w_s   = 1.0f * dexp((float(-kPi) * (gup.a * gup.a)) * (d_i * d_i));
                D2<float> phi_s = phi - kTwoPi * d_i * omega_t.z;
                D2<float> w_f   = filter_weight(fc, w_s);
                D2<V3> xkit  = mult(gp.local, xk);
                D2<float> gk = kernel2(w_f, fc.omf_x, fc.omf_y, phi_s, fc.a_f, comp(xkit, 0),
                                       comp(xkit, 1));
                if (!std::isfinite(gk.v))
                    gk = kernel3(D2<float>(1.0f), omega, D2<float>(phi), gup.a, xk);
                sum = sum + gk;
            } else {
                sum = sum + kernel3(D2<float>(1.0f), omega, D2<float>(phi), gup.a, xk);
            }
        } else {
            // gabor_no_sample: advance the generator as the sampler would
            if (opt.anisotropic == 0) {
                rng();
                rng();
            } else if (opt.anisotropic != 1) {
                rng();
            }
            rng();
        }
    }
    return sum;
}

}  // namespace

GaborUniform gabor_uniform(const GaborOpts& opt)
{
    GaborUniform g;
    float bandwidth = std::min<float>(std::max<float>(opt.bandwidth, 0.01f), 100.0f);
    float two_bw    = exp2f(bandwidth);
    const float SQRT_PI_OVER_LN2 = 2.128934e+00f;
    g.a          = 2.0f * ((two_bw - 1.0f) / (two_bw + 1.0f)) * SQRT_PI_OVER_LN2;
    g.radius     = sqrtf(-logf(0.02f) / float(M_PI)) / g.a;
    g.radius2    = g.radius * g.radius;
    g.radius3    = g.radius2 * g.radius;
    g.radius_inv = 1.0f / g.radius;
    float imp    = std::min<float>(std::max<float>(opt.impulses, 1.0f), 32.0f);
    g.lambda     = imp / (float(1.33333 * M_PI) * g.radius3);
    g.sqrt_lambda_inv = 1.0f / sqrtf(g.lambda);
    g.poisson_g       = expf(-(g.lambda * g.radius3));
    return g;
}

D2<float> gabor(int dim, const D2<float>* p, const GaborOpts& opt)
{
    return gabor_ex(dim, p, opt, 0, nullptr);
}

D2<V3> gabor3(int dim, const D2<float>* p, const GaborOpts& opt, const float* period)
{
    D2<float> r[3];
    for (int seed = 0; seed < 3; ++seed)
        r[seed] = gabor_ex(dim, p, opt, seed, period);
    return D2<V3>(V3(r[0].v, r[1].v, r[2].v), V3(r[0].dx, r[1].dx, r[2].dx),
                  V3(r[0].dy, r[1].dy, r[2].dy));
}

D2<float> gabor_ex(int dim, const D2<float>* p, const GaborOpts& opt, int seed,
                   const float* period)
{
    GaborUniform gup = gabor_uniform(opt);
    const float zero = 0.0f;
    // pgabor of 1 and 2 dimensions pads the period with zeros
    // (gabornoise.cpp:338-352)
    V3 per(period ? period[0] : zero, period && dim > 1 ? period[1] : zero,
           period && dim > 2 ? period[2] : zero);
    D2<V3> P(V3(p[0].v, dim > 1 ? p[1].v : zero, dim > 2 ? p[2].v : zero),
             V3(p[0].dx, dim > 1 ? p[1].dx : zero, dim > 2 ? p[2].dx : zero),
             V3(p[0].dy, dim > 1 ? p[1].dy : zero, dim > 2 ? p[2].dy : zero));
    Params gp;
    gp.omega     = opt.direction;
    gp.do_filter = 0;
    gp.N         = V3(0.0f);
    if (opt.do_filter)
        setup_filter(P, gp);
    if (gp.do_filter == 1 && opt.anisotropic == 1) {
        gp.omega_t = mult(gp.local, gp.omega);
        gp.fc      = filter_consts(gp.filter, gup.a, gp.omega_t.x, gp.omega_t.y);
    }
    // gabor_evaluate + gabor_grid
    D2<V3> xg = P * gup.radius_inv;
    V3 fl(floorf(xg.v.x), floorf(xg.v.y), floorf(xg.v.z));
    D2<V3> xc(xg.v - fl, xg.dx, xg.dy);
    D2<float> sum = D2<float>(0.0f);
    for (int k = -1; k <= 1; ++k)
        for (int j = -1; j <= 1; ++j)
            for (int i = -1; i <= 1; ++i) {
                V3 c((float)i, (float)j, (float)k);
                V3 c_i = fl + c;
                D2<V3> x_c_i(xc.v - c, xc.dx, xc.dy);
                sum = sum + cell(gup, opt, gp, c_i, x_c_i, seed, period ? &per : nullptr);
            }
    D2<float> result   = sum * gup.sqrt_lambda_inv;
    float a3           = gup.a * gup.a * gup.a;
    float gabor_var    = 1.0f / (4.0f * sqrtf(2.0f) * a3);
    float scale        = 1.0f / (3.0f * sqrtf(gabor_var));
    scale *= 0.5f;
    return result * scale;
}

}  // namespace oref
