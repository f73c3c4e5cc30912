// Op-counting scalar types for the oracle (test infrastructure).
//
// The counting build compiles the oracle sources with `float` and
// `uint32_t` replaced by CFloat / CU32, whose operators increment a
// thread-local counter: one count per 32-bit arithmetic, logic, compare,
// convert or math-library operation (an fma counts once; a transcendental
// written as a polynomial counts its polynomial's operations). Integer
// selection logic on plain ints is counted where the oracle marks it with
// OREF_COUNT(k). The totals give the algorithmic 32-bit lane-operations per
// shading point used as the VALU-roofline numerator.
#pragma once
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace ocount {
extern thread_local uint64_t g_ops;
inline void add(uint64_t k) { g_ops += k; }
}  // namespace ocount

#define OREF_COUNT(k) ::ocount::add(k)

struct CU32;
struct CFloat {
    float v;
    CFloat() : v(0.0f) {}
    CFloat(float x) : v(x) {}
    CFloat(double x) : v((float)x) {}
    CFloat(int x) : v((float)x) { ocount::add(1); }
    CFloat(unsigned x) : v((float)x) { ocount::add(1); }
    CFloat(const CU32& x);
    explicit operator int() const { ocount::add(1); return (int)v; }
    explicit operator unsigned() const { ocount::add(1); return (unsigned)v; }
    operator float() const { return v; }
};
#define CF_BIN(op)                                                                          \
    inline CFloat operator op(CFloat a, CFloat b) { ocount::add(1); return CFloat(a.v op b.v); } \
    inline CFloat operator op(CFloat a, float b) { ocount::add(1); return CFloat(a.v op b); }    \
    inline CFloat operator op(float a, CFloat b) { ocount::add(1); return CFloat(a op b.v); }    \
    inline CFloat operator op(CFloat a, double b) { ocount::add(1); return CFloat(a.v op (float)b); } \
    inline CFloat operator op(double a, CFloat b) { ocount::add(1); return CFloat((float)a op b.v); } \
    inline CFloat operator op(CFloat a, int b) { ocount::add(1); return CFloat(a.v op (float)b); }   \
    inline CFloat operator op(int a, CFloat b) { ocount::add(1); return CFloat((float)a op b.v); }
CF_BIN(+)
CF_BIN(-)
CF_BIN(*)
CF_BIN(/)
#undef CF_BIN
#define CF_CMP(op)                                                                     \
    inline bool operator op(CFloat a, CFloat b) { ocount::add(1); return a.v op b.v; } \
    inline bool operator op(CFloat a, float b) { ocount::add(1); return a.v op b; }    \
    inline bool operator op(float a, CFloat b) { ocount::add(1); return a op b.v; }    \
    inline bool operator op(CFloat a, double b) { ocount::add(1); return a.v op (float)b; } \
    inline bool operator op(CFloat a, int b) { ocount::add(1); return a.v op (float)b; }
CF_CMP(<)
CF_CMP(>)
CF_CMP(<=)
CF_CMP(>=)
CF_CMP(==)
CF_CMP(!=)
#undef CF_CMP
inline CFloat operator-(CFloat a) { ocount::add(1); return CFloat(-a.v); }
inline CFloat& operator+=(CFloat& a, CFloat b) { ocount::add(1); a.v += b.v; return a; }
inline CFloat& operator-=(CFloat& a, CFloat b) { ocount::add(1); a.v -= b.v; return a; }
inline CFloat& operator*=(CFloat& a, CFloat b) { ocount::add(1); a.v *= b.v; return a; }
inline CFloat& operator/=(CFloat& a, CFloat b) { ocount::add(1); a.v /= b.v; return a; }

struct CU32 {
    uint32_t v;
    CU32() : v(0) {}
    CU32(uint32_t x) : v(x) {}
    CU32(int x) : v((uint32_t)x) {}
    CU32(unsigned long x) : v((uint32_t)x) {}
    CU32(long x) : v((uint32_t)x) {}
    explicit CU32(const CFloat& f) : v((uint32_t)f.v) { ocount::add(1); }
    operator uint32_t() const { return v; }
};
inline CFloat::CFloat(const CU32& x) : v((float)x.v) { ocount::add(1); }
#define CU_BIN(op)                                                                      \
    inline CU32 operator op(CU32 a, CU32 b) { ocount::add(1); return CU32(a.v op b.v); } \
    inline CU32 operator op(CU32 a, uint32_t b) { ocount::add(1); return CU32(a.v op b); } \
    inline CU32 operator op(uint32_t a, CU32 b) { ocount::add(1); return CU32(a op b.v); } \
    inline CU32 operator op(CU32 a, int b) { ocount::add(1); return CU32(a.v op (uint32_t)b); }
CU_BIN(+)
CU_BIN(-)
CU_BIN(*)
CU_BIN(^)
CU_BIN(&)
CU_BIN(|)
CU_BIN(/)
#undef CU_BIN
inline CU32 operator<<(CU32 a, int k) { ocount::add(1); return CU32(a.v << k); }
inline CU32 operator>>(CU32 a, int k) { ocount::add(1); return CU32(a.v >> k); }
inline CU32& operator+=(CU32& a, CU32 b) { ocount::add(1); a.v += b.v; return a; }
inline CU32& operator-=(CU32& a, CU32 b) { ocount::add(1); a.v -= b.v; return a; }
inline CU32& operator^=(CU32& a, CU32 b) { ocount::add(1); a.v ^= b.v; return a; }
inline CU32& operator*=(CU32& a, CU32 b) { ocount::add(1); a.v *= b.v; return a; }
inline bool operator!(CU32 a) { ocount::add(1); return !a.v; }
// loop control on counted unsigneds (scalar work on the GPU): not counted
inline bool operator<(CU32 a, int b) { return a.v < (uint32_t)b; }
inline bool operator>(CU32 a, int b) { return a.v > (uint32_t)b; }
inline CU32& operator++(CU32& a) { ++a.v; return a; }
inline CU32& operator-=(CU32& a, int b) { a.v -= (uint32_t)b; return a; }

// math library: one op each (hardware instructions on the GPU: floor,
// sqrt, fma, rint, fabs, copysign, min/max) ...
#define CF_UN(fn)                                                    \
    inline CFloat fn(CFloat a) { ocount::add(1); return CFloat(::fn(a.v)); }
CF_UN(floorf)
CF_UN(sqrtf)
CF_UN(fabsf)
CF_UN(rintf)
#undef CF_UN
template<class A, class B, class C> inline CFloat fmaf(A a, B b, C c)
{
    ocount::add(1);
    return CFloat(::fmaf((float)a, (float)b, (float)c));
}
template<class A, class B> inline CFloat copysignf(A a, B b)
{
    ocount::add(1);
    return CFloat(::copysignf((float)a, (float)b));
}
// ... host-only library calls (table builds, uniform setup) are not on the
// per-point path and count one op each.
inline CFloat expm1f(CFloat a) { ocount::add(1); return CFloat(::expm1f(a.v)); }
inline CFloat expf(CFloat a) { ocount::add(1); return CFloat(::expf(a.v)); }
inline CFloat logf(CFloat a) { ocount::add(1); return CFloat(::logf(a.v)); }
inline CFloat exp2f(CFloat a) { ocount::add(1); return CFloat(::exp2f(a.v)); }
template<class A, class B> inline CFloat powf(A a, B b)
{
    ocount::add(1);
    return CFloat(::powf((float)a, (float)b));
}
namespace std {
inline CFloat max(CFloat a, CFloat b) { ocount::add(1); return a.v < b.v ? b : a; }
inline CFloat min(CFloat a, CFloat b) { ocount::add(1); return b.v < a.v ? b : a; }
inline bool isfinite(CFloat a) { ocount::add(1); return std::isfinite(a.v); }
}  // namespace std
