// 4x4 matrices and the triple transforms OSL applies with them, for host
// code (the named-transform tables) and device code (emitted kernels).
//
// Layout and convention are Imath's: row vectors, p' = p M, element
// x[i][j] at m[4 i + j], translation in m[12..14]. The operations restate,
// one IEEE operation per source operator in the same order, OSL's own
// helpers the batched shadeops call:
//   multiplyMatrixByMatrix    src/include/OSL/Imathx/Imathx.h:405-460
//   test_if_affine            Imathx.h:130-136
//   affineInverse             Imathx.h:138-230
//   nonAffineInverse          Imathx.h:244-350 (Gauss-Jordan, partial pivot)
//   inlinedTransposed         Imathx.h:382-400
//   multDirMatrix             Imathx.h:47-56 (and dual_vec.h:389-396)
//   robust_multVecMatrix      src/include/OSL/dual_vec.h:336-385
// and the choice of inverse the batched transform of normals makes
// (src/liboslexec/wide/wide_opmatrix.cpp:1410-1445).
#pragma once
#include "oslmi_types.h"

namespace oslmi {

struct M44 {
    float m[16];
};

OSLMI_HD M44 m44_identity()
{
    M44 r;
    for (int k = 0; k < 16; ++k)
        r.m[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    return r;
}

// c = a b, each entry summed left to right over k
OSLMI_HD M44 m44_mul(const M44& a, const M44& b)
{
    M44 c;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            c.m[4 * i + j] = a.m[4 * i + 0] * b.m[0 + j] + a.m[4 * i + 1] * b.m[4 + j]
                             + a.m[4 * i + 2] * b.m[8 + j] + a.m[4 * i + 3] * b.m[12 + j];
    return c;
}

OSLMI_HD bool m44_is_affine(const M44& a)
{
    return a.m[3] == 0.0f && a.m[7] == 0.0f && a.m[11] == 0.0f && a.m[15] == 1.0f;
}

OSLMI_HD float m44_fabs(float v) { return v < 0.0f ? -v : v; }

// Inverse of an affine matrix by cofactors of the 3x3 block; identity when
// the determinant is too small to divide by safely.
OSLMI_HD M44 m44_affine_inverse(const M44& a)
{
    const float* x = a.m;
    M44 s;
    float* t = s.m;
    t[0]  = x[5] * x[10] - x[9] * x[6];
    t[1]  = x[9] * x[2] - x[1] * x[10];
    t[2]  = x[1] * x[6] - x[5] * x[2];
    t[3]  = 0.0f;
    t[4]  = x[8] * x[6] - x[4] * x[10];
    t[5]  = x[0] * x[10] - x[8] * x[2];
    t[6]  = x[4] * x[2] - x[0] * x[6];
    t[7]  = 0.0f;
    t[8]  = x[4] * x[9] - x[8] * x[5];
    t[9]  = x[8] * x[1] - x[0] * x[9];
    t[10] = x[0] * x[5] - x[4] * x[1];
    t[11] = 0.0f;
    t[12] = t[13] = t[14] = 0.0f;
    t[15] = 1.0f;
    const float r     = x[0] * t[0] + x[1] * t[4] + x[2] * t[8];
    const float abs_r = m44_fabs(r);
    bool singular     = false;
    if (abs_r < 1.0f) {
        const float mr = abs_r / 1.17549435e-38f;  // FLT_MIN
        const int idx[9] = { 0, 1, 2, 4, 5, 6, 8, 9, 10 };
        for (int k = 0; k < 9; ++k)
            singular = singular | (mr <= m44_fabs(t[idx[k]]));
    }
    const int idx[9] = { 0, 1, 2, 4, 5, 6, 8, 9, 10 };
    for (int k = 0; k < 9; ++k)
        t[idx[k]] /= r;
    t[12] = -x[12] * t[0] - x[13] * t[4] - x[14] * t[8];
    t[13] = -x[12] * t[1] - x[13] * t[5] - x[14] * t[9];
    t[14] = -x[12] * t[2] - x[13] * t[6] - x[14] * t[10];
    return singular ? m44_identity() : s;
}

// General inverse by Gauss-Jordan elimination with partial pivoting;
// identity for a singular matrix.
OSLMI_HD M44 m44_nonaffine_inverse(const M44& a)
{
    M44 t = a;
    M44 s = m44_identity();
    for (int i = 0; i < 3; i++) {
        int pivot       = i;
        float pivotsize = t.m[4 * i + i];
        if (pivotsize < 0)
            pivotsize = -pivotsize;
        for (int j = i + 1; j < 4; j++) {
            float tmp = t.m[4 * j + i];
            if (tmp < 0)
                tmp = -tmp;
            if (tmp > pivotsize) {
                pivot     = j;
                pivotsize = tmp;
            }
        }
        if (pivotsize == 0)
            return m44_identity();
        if (pivot != i) {
            for (int j = 0; j < 4; j++) {
                float tmp            = t.m[4 * i + j];
                t.m[4 * i + j]       = t.m[4 * pivot + j];
                t.m[4 * pivot + j]   = tmp;
                tmp                  = s.m[4 * i + j];
                s.m[4 * i + j]       = s.m[4 * pivot + j];
                s.m[4 * pivot + j]   = tmp;
            }
        }
        for (int j = i + 1; j < 4; j++) {
            const float f = t.m[4 * j + i] / t.m[4 * i + i];
            for (int k = 0; k < 4; k++) {
                t.m[4 * j + k] -= f * t.m[4 * i + k];
                s.m[4 * j + k] -= f * s.m[4 * i + k];
            }
        }
    }
    for (int i = 3; i >= 0; --i) {
        float f = t.m[4 * i + i];
        if (f == 0)
            return m44_identity();
        for (int j = 0; j < 4; j++) {
            t.m[4 * i + j] /= f;
            s.m[4 * i + j] /= f;
        }
        for (int j = 0; j < i; j++) {
            f = t.m[4 * j + i];
            for (int k = 0; k < 4; k++) {
                t.m[4 * j + k] -= f * t.m[4 * i + k];
                s.m[4 * j + k] -= f * s.m[4 * i + k];
            }
        }
    }
    return s;
}

OSLMI_HD M44 m44_inverse(const M44& a)
{
    return m44_is_affine(a) ? m44_affine_inverse(a) : m44_nonaffine_inverse(a);
}

OSLMI_HD M44 m44_transposed(const M44& a)
{
    M44 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[4 * i + j] = a.m[4 * j + i];
    return r;
}

// point: homogeneous divide, (0,0,0) where w is exactly zero
OSLMI_HD V3 xform_point(const M44& M, V3 p)
{
    const float* m = M.m;
    const float a  = p.x * m[0] + p.y * m[4] + p.z * m[8] + m[12];
    const float b  = p.x * m[1] + p.y * m[5] + p.z * m[9] + m[13];
    const float c  = p.x * m[2] + p.y * m[6] + p.z * m[10] + m[14];
    const float w  = p.x * m[3] + p.y * m[7] + p.z * m[11] + m[15];
    if (w != 0.0f)
        return V3(a / w, b / w, c / w);
    return V3(0.0f, 0.0f, 0.0f);
}
// vector: the upper 3x3 block only
OSLMI_HD V3 xform_vector(const M44& M, V3 p)
{
    const float* m = M.m;
    return V3(p.x * m[0] + p.y * m[4] + p.z * m[8], p.x * m[1] + p.y * m[5] + p.z * m[9],
              p.x * m[2] + p.y * m[6] + p.z * m[10]);
}
// normal: by the transposed inverse
OSLMI_HD V3 xform_normal(const M44& M, V3 n)
{
    return xform_vector(m44_transposed(m44_inverse(M)), n);
}

// A slot of the shading system's transform table: the from-to matrix,
// the transposed inverse used for normals, and whether both spaces were
// known to the renderer.
OSLMI_HD const M44& xf_matrix(const float* slot) { return *reinterpret_cast<const M44*>(slot); }
OSLMI_HD const M44& xf_normal_matrix(const float* slot)
{
    return *reinterpret_cast<const M44*>(slot + 16);
}
OSLMI_HD bool xf_ok(const float* slot) { return slot[32] != 0.0f; }

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
// the same with derivatives (each component a Dual2)
OSLMI_DEV Dv xform_point(const M44& M, const Dv& p)
{
    const float* m = M.m;
    const Df x = comp(p, 0), y = comp(p, 1), z = comp(p, 2);
    const Df a = x * m[0] + y * m[4] + z * m[8] + m[12];
    const Df b = x * m[1] + y * m[5] + z * m[9] + m[13];
    const Df c = x * m[2] + y * m[6] + z * m[10] + m[14];
    const Df w = x * m[3] + y * m[7] + z * m[11] + m[15];
    if (w.v != 0.0f)
        return make_dv(a / w, b / w, c / w);
    return Dv(V3(0.0f));
}
OSLMI_DEV Dv xform_vector(const M44& M, const Dv& p)
{
    return Dv(xform_vector(M, p.v), xform_vector(M, p.dx), xform_vector(M, p.dy));
}
OSLMI_DEV Dv xform_normal(const M44& M, const Dv& n)
{
    return xform_vector(m44_transposed(m44_inverse(M)), n);
}
#endif

}  // namespace oslmi
