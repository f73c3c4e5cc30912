// CPU oracle: 4x4 matrices and triple transforms.
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
//
// Restates, in binary32 with one IEEE operation per source operator, the
// helpers OSL's batched matrix shadeops use (src/include/OSL/Imathx/
// Imathx.h): multiplyMatrixByMatrix (:405-460), test_if_affine (:130),
// affineInverse (:138-230), nonAffineInverse (:244-350), inlinedTransposed
// (:382), multDirMatrix (:47); and robust_multVecMatrix (dual_vec.h:336).
// Row-vector convention: p' = p M, translation in row 3.
#pragma once
#include <cfloat>
#include <utility>

#include "oracle_core.h"

namespace oref {

struct Mat4 {
    float x[4][4];
};

inline Mat4 mat4_identity()
{
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.x[i][j] = i == j ? 1.0f : 0.0f;
    return r;
}
inline Mat4 mat4_from(const float* f)
{
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.x[i][j] = f[4 * i + j];
    return r;
}

// (a b)[i][j] = a[i][0] b[0][j] + a[i][1] b[1][j] + a[i][2] b[2][j] + a[i][3] b[3][j]
inline Mat4 mat4_mul(const Mat4& a, const Mat4& b)
{
    Mat4 c;
    for (int i = 0; i < 4; ++i) {
        const float ai0 = a.x[i][0], ai1 = a.x[i][1], ai2 = a.x[i][2], ai3 = a.x[i][3];
        for (int j = 0; j < 4; ++j)
            c.x[i][j] = ai0 * b.x[0][j] + ai1 * b.x[1][j] + ai2 * b.x[2][j] + ai3 * b.x[3][j];
    }
    return c;
}

inline Mat4 mat4_transposed(const Mat4& a)
{
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.x[j][i] = a.x[i][j];
    return r;
}

inline bool mat4_affine(const Mat4& m)
{
    return m.x[0][3] == 0.0f && m.x[1][3] == 0.0f && m.x[2][3] == 0.0f && m.x[3][3] == 1.0f;
}

// Cofactor inverse of the 3x3 block plus the translation row; identity when
// the determinant is too small to divide the cofactors by.
inline Mat4 mat4_affine_inverse(const Mat4& m)
{
    Mat4 s = mat4_identity();
    const float(*x)[4] = m.x;
    s.x[0][0] = x[1][1] * x[2][2] - x[2][1] * x[1][2];
    s.x[0][1] = x[2][1] * x[0][2] - x[0][1] * x[2][2];
    s.x[0][2] = x[0][1] * x[1][2] - x[1][1] * x[0][2];
    s.x[1][0] = x[2][0] * x[1][2] - x[1][0] * x[2][2];
    s.x[1][1] = x[0][0] * x[2][2] - x[2][0] * x[0][2];
    s.x[1][2] = x[1][0] * x[0][2] - x[0][0] * x[1][2];
    s.x[2][0] = x[1][0] * x[2][1] - x[2][0] * x[1][1];
    s.x[2][1] = x[2][0] * x[0][1] - x[0][0] * x[2][1];
    s.x[2][2] = x[0][0] * x[1][1] - x[1][0] * x[0][1];
    const float r  = x[0][0] * s.x[0][0] + x[0][1] * s.x[1][0] + x[0][2] * s.x[2][0];
    const float ar = std::fabs(r);
    bool bad       = false;
    if (ar < 1.0f) {
        const float mr = ar / FLT_MIN;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (mr <= std::fabs(s.x[i][j]))
                    bad = true;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            s.x[i][j] /= r;
    for (int j = 0; j < 3; ++j)
        s.x[3][j] = -x[3][0] * s.x[0][j] - x[3][1] * s.x[1][j] - x[3][2] * s.x[2][j];
    return bad ? mat4_identity() : s;
}

// Gauss-Jordan with partial pivoting on the first three columns, then back
// substitution; identity when a pivot is zero.
inline Mat4 mat4_nonaffine_inverse(const Mat4& m)
{
    Mat4 t = m, s = mat4_identity();
    for (int i = 0; i < 3; i++) {
        int pivot = i;
        float big = std::fabs(t.x[i][i]);
        for (int j = i + 1; j < 4; j++) {
            const float v = std::fabs(t.x[j][i]);
            if (v > big) {
                pivot = j;
                big   = v;
            }
        }
        if (big == 0.0f)
            return mat4_identity();
        if (pivot != i)
            for (int j = 0; j < 4; j++) {
                std::swap(t.x[i][j], t.x[pivot][j]);
                std::swap(s.x[i][j], s.x[pivot][j]);
            }
        for (int j = i + 1; j < 4; j++) {
            const float f = t.x[j][i] / t.x[i][i];
            for (int k = 0; k < 4; k++) {
                t.x[j][k] -= f * t.x[i][k];
                s.x[j][k] -= f * s.x[i][k];
            }
        }
    }
    for (int i = 3; i >= 0; --i) {
        const float d = t.x[i][i];
        if (d == 0.0f)
            return mat4_identity();
        for (int j = 0; j < 4; j++) {
            t.x[i][j] /= d;
            s.x[i][j] /= d;
        }
        for (int j = 0; j < i; j++) {
            const float f = t.x[j][i];
            for (int k = 0; k < 4; k++) {
                t.x[j][k] -= f * t.x[i][k];
                s.x[j][k] -= f * s.x[i][k];
            }
        }
    }
    return s;
}

inline Mat4 mat4_inverse(const Mat4& m)
{
    return mat4_affine(m) ? mat4_affine_inverse(m) : mat4_nonaffine_inverse(m);
}

inline V3 mat4_point(const Mat4& M, V3 p)
{
    const float a = p.x * M.x[0][0] + p.y * M.x[1][0] + p.z * M.x[2][0] + M.x[3][0];
    const float b = p.x * M.x[0][1] + p.y * M.x[1][1] + p.z * M.x[2][1] + M.x[3][1];
    const float c = p.x * M.x[0][2] + p.y * M.x[1][2] + p.z * M.x[2][2] + M.x[3][2];
    const float w = p.x * M.x[0][3] + p.y * M.x[1][3] + p.z * M.x[2][3] + M.x[3][3];
    if (w == 0.0f)
        return V3(0.0f, 0.0f, 0.0f);
    return V3(a / w, b / w, c / w);
}
inline V3 mat4_dir(const Mat4& M, V3 v)
{
    return V3(v.x * M.x[0][0] + v.y * M.x[1][0] + v.z * M.x[2][0],
              v.x * M.x[0][1] + v.y * M.x[1][1] + v.z * M.x[2][1],
              v.x * M.x[0][2] + v.y * M.x[1][2] + v.z * M.x[2][2]);
}
inline V3 mat4_normal(const Mat4& M, V3 n) { return mat4_dir(mat4_transposed(mat4_inverse(M)), n); }

}  // namespace oref
