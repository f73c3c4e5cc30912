// CPU oracle for the batched shading path: shared scalar types.
//
// TEST INFRASTRUCTURE ONLY. This directory restates the reference's
// algorithms on the CPU so that tests, the smoke check and the CPU-baseline
// leg of the benchmark can check the HIP kernels against it. Nothing in the
// product (openshadinglanguage_amd/) links, loads or calls it.
//
// Floating point policy: every file here is compiled with -ffp-contract=off
// and no fast-math, so each +,-,*,/ is one IEEE-754 binary32 operation in
// source order. Fused multiply-adds appear only where written as fmaf().
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

// Op-count marks for integer selection logic (see count/counted.h); no-ops
// in the normal build.
#ifndef OREF_COUNT
#define OREF_COUNT(k) ((void)0)
#endif

namespace oref {

struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit V3(float a) : x(a), y(a), z(a) {}
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3 operator*(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator/(V3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
// Imath Vec3::dot: x*x' + y*y' + z*z' evaluated left to right.
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b)
{
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
              a.x * b.y - a.y * b.x);
}

// Value plus two partial derivatives (dx, dy): src/include/OSL/dual.h:228.
// Operator bodies follow dual.h:420-600 (elementwise +/-, product rule for
// *, quotient rule with a shared reciprocal for /).
template<class T> struct D2 {
    T v, dx, dy;
    D2() : v(), dx(), dy() {}
    D2(const T& a) : v(a), dx(T(0)), dy(T(0)) {}
    D2(const T& a, const T& b, const T& c) : v(a), dx(b), dy(c) {}
};
template<class T> inline D2<T> operator+(const D2<T>& a, const D2<T>& b)
{
    return D2<T>(a.v + b.v, a.dx + b.dx, a.dy + b.dy);
}
template<class T> inline D2<T> operator-(const D2<T>& a, const D2<T>& b)
{
    return D2<T>(a.v - b.v, a.dx - b.dx, a.dy - b.dy);
}
template<class T> inline D2<T> operator-(const D2<T>& a)
{
    return D2<T>(-a.v, -a.dx, -a.dy);
}
template<class T> inline D2<T> operator+(const D2<T>& a, float b)
{
    return D2<T>(a.v + b, a.dx, a.dy);
}
template<class T> inline D2<T> operator-(const D2<T>& a, float b)
{
    return D2<T>(a.v - b, a.dx, a.dy);
}
// scalar - dual: dual.h:460-470 (partials negated).
inline D2<float> operator-(float a, const D2<float>& b)
{
    return D2<float>(a - b.v, -b.dx, -b.dy);
}
inline D2<float> operator*(const D2<float>& a, const D2<float>& b)
{
    return D2<float>(a.v * b.v, a.v * b.dx + a.dx * b.v,
                     a.v * b.dy + a.dy * b.v);
}
inline D2<V3> operator*(const D2<V3>& a, const D2<float>& b)
{
    return D2<V3>(a.v * b.v, a.v * b.dx + a.dx * b.v,
                  a.v * b.dy + a.dy * b.v);
}
inline D2<V3> operator*(const D2<float>& a, const D2<V3>& b)
{
    return D2<V3>(a.v * b.v, a.v * b.dx + a.dx * b.v,
                  a.v * b.dy + a.dy * b.v);
}
template<class T> inline D2<T> operator*(const D2<T>& a, float s)
{
    return D2<T>(a.v * s, a.dx * s, a.dy * s);
}
template<class T> inline D2<T> operator*(float s, const D2<T>& a)
{
    return D2<T>(a.v * s, a.dx * s, a.dy * s);
}
inline D2<float> operator/(const D2<float>& a, const D2<float>& b)
{
    float binv = 1.0f / b.v;
    float ab   = a.v / b.v;
    return D2<float>(ab, binv * (a.dx - ab * b.dx), binv * (a.dy - ab * b.dy));
}
// scalar / dual and dual / scalar (dual.h:592-618)
inline D2<float> operator/(float a, const D2<float>& b)
{
    float binv = 1.0f / b.v;
    float ab   = a / b.v;
    return D2<float>(ab, binv * (-ab * b.dx), binv * (-ab * b.dy));
}
inline D2<float> operator/(const D2<float>& a, float b)
{
    float binv = 1.0f / b;
    return D2<float>(a.v / b, binv * a.dx, binv * a.dy);
}

#ifdef OREF_F64
// FP64 evaluation build: hashes still see the binary32 value of a
// coordinate (oref_f32 is the real 4-byte float there), so lattice and
// cell choices match the binary32 path wherever the coordinate rounds alike.
inline uint32_t f2u(float f)
{
    const oref_f32 g = (oref_f32)f;
    uint32_t u;
    std::memcpy(&u, &g, 4);
    return u;
}
inline float u2f(uint32_t u)
{
    oref_f32 g;
    std::memcpy(&g, &u, 4);
    return g;
}
#else
inline uint32_t f2u(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float u2f(uint32_t u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
#endif

}  // namespace oref
