// b2g: MI355X-native B2world::step -- scalar math shared by the kernels.
//
// Every operation keeps the operand order of the reference (box2d-rs src/b2_math.rs) and the
// library is built with -ffp-contract=off, so products and sums round exactly as the
// reference's f32 code does. The header compiles as HIP device code and as plain C++ (the
// B2G_HD macro), which lets the host test build run the same collision code on the CPU.
//
// The vector types are templates over the scalar: float for the simulation kernels, double for
// host checks, and the forward-mode Dual of b2g_dual.h for the adjoint's local Jacobians. Generic
// code writes its constants as T(double literal), which for T = float is the same value as the
// float literal of the reference.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define B2G_HD __host__ __device__ inline
#else
#define B2G_HD inline
#endif

namespace b2g {

// src/b2_common.rs:17-91, src/b2_settings.rs:7-11 (f32 values)
constexpr float kMaxFloat = 3.40282347e+38f;
constexpr float kEpsilon = 1.1920929e-07f;
constexpr float kPi = 3.14159265358979323846f;
constexpr float kAabbExtension = 0.1f;
constexpr float kAabbMultiplier = 4.0f;
constexpr float kLinearSlop = 0.005f;
// sleep thresholds (src/b2_common.rs:85-91; length units per metre = 1)
constexpr float kTimeToSleep = 0.5f;
constexpr float kLinearSleepTolerance = 0.01f;
constexpr float kAngularSleepTolerance = 2.0f / 180.0f * kPi;
constexpr float kPolygonRadius = 2.0f * 0.005f;
constexpr float kMaxLinearCorrection = 0.2f;
constexpr float kMaxTranslation = 2.0f;
constexpr float kMaxRotation = 0.5f * kPi;
constexpr float kBaumgarte = 0.2f;
// the same settings as exact decimal values, converted to the generic scalar T at use
constexpr double kLinearSlopD = 0.005;
constexpr double kMaxLinearCorrectionD = 0.2;
constexpr double kBaumgarteD = 0.2;
constexpr double kMaxTranslationD = 2.0;
constexpr double kPiD = 3.14159265358979323846;

// Generic scalar helpers: float/double use the C library; other scalar types (Dual) overload
// tsqrt/tsin/tcos/primal in their own header.
B2G_HD float tsqrt(float x) { return sqrtf(x); }
B2G_HD double tsqrt(double x) { return sqrt(x); }
B2G_HD float primal(float x) { return x; }
B2G_HD double primal(double x) { return x; }

template <class T>
struct TV2 {
  T x, y;
};
using V2 = TV2<float>;

template <class T>
B2G_HD TV2<T> v2(T x, T y) { TV2<T> r; r.x = x; r.y = y; return r; }
// converts a vector of another scalar type (a float shape vertex into the generic scalar)
template <class T, class U>
B2G_HD TV2<T> tv(TV2<U> a) { return v2(T(a.x), T(a.y)); }
template <class T>
B2G_HD TV2<T> operator+(TV2<T> a, TV2<T> b) { return v2(a.x + b.x, a.y + b.y); }
template <class T>
B2G_HD TV2<T> operator-(TV2<T> a, TV2<T> b) { return v2(a.x - b.x, a.y - b.y); }
template <class T>
B2G_HD TV2<T> operator-(TV2<T> a) { return v2(-a.x, -a.y); }
template <class T>
B2G_HD TV2<T> operator*(T s, TV2<T> a) { return v2(s * a.x, s * a.y); }
template <class T>
B2G_HD T dot(TV2<T> a, TV2<T> b) { return a.x * b.x + a.y * b.y; }
template <class T>
B2G_HD T cross(TV2<T> a, TV2<T> b) { return a.x * b.y - a.y * b.x; }
template <class T>
B2G_HD TV2<T> cross(TV2<T> a, T s) { return v2(s * a.y, -s * a.x); }
template <class T>
B2G_HD TV2<T> cross(T s, TV2<T> a) { return v2(-s * a.y, s * a.x); }
template <class T>
B2G_HD T len(TV2<T> a) { return tsqrt(a.x * a.x + a.y * a.y); }
template <class T>
B2G_HD T dist_sq(TV2<T> a, TV2<T> b) { TV2<T> c = a - b; return dot(c, c); }
// b2_min/b2_max/b2_clamp exactly as src/b2_math.rs:704-731
template <class T>
B2G_HD T bmin(T a, T b) { return a < b ? a : b; }
template <class T>
B2G_HD T bmax(T a, T b) { return a > b ? a : b; }
template <class T>
B2G_HD T bclamp(T a, T lo, T hi) { return bmax(lo, bmin(a, hi)); }
template <class T>
B2G_HD TV2<T> vmin(TV2<T> a, TV2<T> b) { return v2(bmin(a.x, b.x), bmin(a.y, b.y)); }
template <class T>
B2G_HD TV2<T> vmax(TV2<T> a, TV2<T> b) { return v2(bmax(a.x, b.x), bmax(a.y, b.y)); }

// B2vec2::normalize (src/b2_math.rs:82-92)
template <class T>
B2G_HD T normalize(TV2<T>* v) {
  T l = len(*v);
  if (l < T(kEpsilon)) return T(0.0);
  T inv = T(1.0) / l;
  v->x *= inv;
  v->y *= inv;
  return l;
}

template <class T>
struct TRot {
  T s, c;
};
using Rot = TRot<float>;
// B2Rot::set (src/b2_math.rs:372-376) is f32::sin / f32::cos, i.e. glibc sinf/cosf, which
// evaluate in double precision and round once to float. The device does the same: double
// sin/cos, one rounding; this matches glibc except when the double result lies within a
// hair of a float rounding boundary.
B2G_HD Rot rot(float angle) {
  Rot q;
#if defined(__HIP_DEVICE_COMPILE__)
  double s, c;
  sincos(static_cast<double>(angle), &s, &c);
  q.s = static_cast<float>(s);
  q.c = static_cast<float>(c);
#else
  q.s = sinf(angle);
  q.c = cosf(angle);
#endif
  return q;
}
B2G_HD TRot<double> rot(double angle) {
  TRot<double> q;
  q.s = sin(angle);
  q.c = cos(angle);
  return q;
}

template <class T>
struct TXf {
  TV2<T> p;
  TRot<T> q;
};
using Xf = TXf<float>;

// src/b2_math.rs:611-680
template <class T>
B2G_HD TV2<T> mul(TRot<T> q, TV2<T> v) { return v2(q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y); }
template <class T>
B2G_HD TV2<T> mulT(TRot<T> q, TV2<T> v) { return v2(q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y); }
template <class T>
B2G_HD TRot<T> mulT(TRot<T> q, TRot<T> r) {
  TRot<T> o;
  o.s = q.c * r.s - q.s * r.c;
  o.c = q.c * r.c + q.s * r.s;
  return o;
}
template <class T>
B2G_HD TV2<T> mul(TXf<T> t, TV2<T> v) {
  T x = (t.q.c * v.x - t.q.s * v.y) + t.p.x;
  T y = (t.q.s * v.x + t.q.c * v.y) + t.p.y;
  return v2(x, y);
}
template <class T>
B2G_HD TV2<T> mulT(TXf<T> t, TV2<T> v) {
  T px = v.x - t.p.x, py = v.y - t.p.y;
  return v2(t.q.c * px + t.q.s * py, -t.q.s * px + t.q.c * py);
}
template <class T>
B2G_HD TXf<T> mulT(TXf<T> a, TXf<T> b) {
  TXf<T> c;
  c.q = mulT(a.q, b.q);
  c.p = mulT(a.q, b.p - a.p);
  return c;
}

// AABB helpers (src/b2_collision.rs:201-256, 355-368)
struct Aabb {
  V2 lo, hi;
};
B2G_HD bool overlap(Aabb a, Aabb b) {
  V2 d1 = b.lo - a.hi, d2 = a.lo - b.hi;
  if (d1.x > 0.0f || d1.y > 0.0f) return false;
  if (d2.x > 0.0f || d2.y > 0.0f) return false;
  return true;
}
B2G_HD bool contains(Aabb outer, Aabb in) {
  bool r = true;
  r = r && outer.lo.x <= in.lo.x;
  r = r && outer.lo.y <= in.lo.y;
  r = r && in.hi.x <= outer.hi.x;
  r = r && in.hi.y <= outer.hi.y;
  return r;
}
B2G_HD Aabb combine(Aabb a, Aabb b) { Aabb r; r.lo = vmin(a.lo, b.lo); r.hi = vmax(a.hi, b.hi); return r; }
B2G_HD V2 center(Aabb a) { return 0.5f * (a.lo + a.hi); }

}  // namespace b2g
