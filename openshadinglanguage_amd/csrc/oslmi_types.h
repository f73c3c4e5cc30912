// Device value types for the CDNA4 shadeop library: 3-vectors and Dual2
// (value + x/y screen derivatives).
//
// The library is compiled with -ffp-contract=off: each arithmetic operator
// below is one IEEE binary32 operation, in the order written. Dual2
// arithmetic follows the product/quotient rules of the reference's
// src/include/OSL/dual.h:420-600, so values and derivatives agree with the
// reference's evaluation order.
#pragma once

#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>

#include <cstdint>
#else
// hiprtc provides the fixed-width integer types in an internal namespace
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
typedef __hip_internal::uint64_t uint64_t;
#endif

#ifndef OSLMI_DEV
#define OSLMI_DEV __device__ __forceinline__
#endif
// Host and device: code the library also runs on the CPU (table bakes).
#ifndef OSLMI_HD
#define OSLMI_HD __host__ __device__ __forceinline__
#endif

namespace oslmi {

struct V3 {
    float x, y, z;
    OSLMI_HD V3() : x(0.f), y(0.f), z(0.f) {}
    OSLMI_HD V3(float a, float b, float c) : x(a), y(b), z(c) {}
    OSLMI_HD explicit V3(float a) : x(a), y(a), z(a) {}
};
OSLMI_HD V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
OSLMI_HD V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
OSLMI_HD V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
OSLMI_HD V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
OSLMI_HD V3 operator*(float s, V3 a) { return V3(s * a.x, s * a.y, s * a.z); }
OSLMI_HD V3 operator*(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
OSLMI_HD V3 operator/(V3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
OSLMI_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
OSLMI_HD V3 cross(V3 a, V3 b)
{
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

template<class T> struct D2 {
    T v, dx, dy;
    OSLMI_DEV D2() : v(), dx(), dy() {}
    OSLMI_DEV D2(const T& a) : v(a), dx(T(0.f)), dy(T(0.f)) {}
    OSLMI_DEV D2(const T& a, const T& b, const T& c) : v(a), dx(b), dy(c) {}
};
typedef D2<float> Df;
typedef D2<V3> Dv;

template<class T> OSLMI_DEV D2<T> operator+(const D2<T>& a, const D2<T>& b)
{
    return D2<T>(a.v + b.v, a.dx + b.dx, a.dy + b.dy);
}
template<class T> OSLMI_DEV D2<T> operator-(const D2<T>& a, const D2<T>& b)
{
    return D2<T>(a.v - b.v, a.dx - b.dx, a.dy - b.dy);
}
template<class T> OSLMI_DEV D2<T> operator-(const D2<T>& a) { return D2<T>(-a.v, -a.dx, -a.dy); }
template<class T> OSLMI_DEV D2<T> operator+(const D2<T>& a, float b) { return D2<T>(a.v + b, a.dx, a.dy); }
template<class T> OSLMI_DEV D2<T> operator-(const D2<T>& a, float b) { return D2<T>(a.v - b, a.dx, a.dy); }
OSLMI_DEV Df operator-(float a, const Df& b) { return Df(a - b.v, -b.dx, -b.dy); }
OSLMI_DEV Df operator+(float a, const Df& b) { return Df(a + b.v, b.dx, b.dy); }
OSLMI_DEV Df operator*(const Df& a, const Df& b)
{
    return Df(a.v * b.v, a.v * b.dx + a.dx * b.v, a.v * b.dy + a.dy * b.v);
}
OSLMI_DEV Dv operator*(const Dv& a, const Df& b)
{
    return Dv(a.v * b.v, a.v * b.dx + a.dx * b.v, a.v * b.dy + a.dy * b.v);
}
OSLMI_DEV Dv operator*(const Df& a, const Dv& b)
{
    return Dv(a.v * b.v, a.v * b.dx + a.dx * b.v, a.v * b.dy + a.dy * b.v);
}
template<class T> OSLMI_DEV D2<T> operator*(const D2<T>& a, float s)
{
    return D2<T>(a.v * s, a.dx * s, a.dy * s);
}
template<class T> OSLMI_DEV D2<T> operator*(float s, const D2<T>& a)
{
    return D2<T>(a.v * s, a.dx * s, a.dy * s);
}
OSLMI_DEV Df operator/(const Df& a, const Df& b)
{
    float binv = 1.0f / b.v;
    float ab   = a.v / b.v;
    return Df(ab, binv * (a.dx - ab * b.dx), binv * (a.dy - ab * b.dy));
}
OSLMI_DEV Df operator/(float a, const Df& b)
{
    float binv = 1.0f / b.v;
    float ab   = a / b.v;
    return Df(ab, binv * (-ab * b.dx), binv * (-ab * b.dy));
}
OSLMI_DEV Df operator/(const Df& a, float b)
{
    float binv = 1.0f / b;
    return Df(a.v / b, binv * a.dx, binv * a.dy);
}
OSLMI_DEV Df comp(const Dv& a, int i)
{
    return i == 0 ? Df(a.v.x, a.dx.x, a.dy.x)
                  : (i == 1 ? Df(a.v.y, a.dx.y, a.dy.y) : Df(a.v.z, a.dx.z, a.dy.z));
}
OSLMI_DEV Dv make_dv(const Df& a, const Df& b, const Df& c)
{
    return Dv(V3(a.v, b.v, c.v), V3(a.dx, b.dx, c.dx), V3(a.dy, b.dy, c.dy));
}

}  // namespace oslmi
