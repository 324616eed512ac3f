// b2g: MI355X-native B2world::step -- scalar math shared by the kernels.
//
// Every operation keeps the operand order of the reference (box2d-rs src/b2_math.rs) and the
// library is built with -ffp-contract=off, so products and sums round exactly as the
// reference's f32 code does. The header compiles as HIP device code and as plain C++ (the
// B2G_HD macro), which lets the host test build run the same collision code on the CPU.
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
constexpr float kPolygonRadius = 2.0f * 0.005f;
constexpr float kMaxLinearCorrection = 0.2f;
constexpr float kMaxTranslation = 2.0f;
constexpr float kMaxRotation = 0.5f * kPi;
constexpr float kBaumgarte = 0.2f;

struct V2 {
  float x, y;
};
B2G_HD V2 v2(float x, float y) { V2 r; r.x = x; r.y = y; return r; }
B2G_HD V2 operator+(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
B2G_HD V2 operator-(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }
B2G_HD V2 operator-(V2 a) { return v2(-a.x, -a.y); }
B2G_HD V2 operator*(float s, V2 a) { return v2(s * a.x, s * a.y); }
B2G_HD float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
B2G_HD float cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
B2G_HD V2 cross(V2 a, float s) { return v2(s * a.y, -s * a.x); }
B2G_HD V2 cross(float s, V2 a) { return v2(-s * a.y, s * a.x); }
B2G_HD float len(V2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
B2G_HD float dist_sq(V2 a, V2 b) { V2 c = a - b; return dot(c, c); }
// b2_min/b2_max/b2_clamp exactly as src/b2_math.rs:704-731
B2G_HD float bmin(float a, float b) { return a < b ? a : b; }
B2G_HD float bmax(float a, float b) { return a > b ? a : b; }
B2G_HD float bclamp(float a, float lo, float hi) { return bmax(lo, bmin(a, hi)); }
B2G_HD V2 vmin(V2 a, V2 b) { return v2(bmin(a.x, b.x), bmin(a.y, b.y)); }
B2G_HD V2 vmax(V2 a, V2 b) { return v2(bmax(a.x, b.x), bmax(a.y, b.y)); }

// B2vec2::normalize (src/b2_math.rs:82-92)
B2G_HD float normalize(V2* v) {
  float l = len(*v);
  if (l < kEpsilon) return 0.0f;
  float inv = 1.0f / l;
  v->x *= inv;
  v->y *= inv;
  return l;
}

struct Rot {
  float s, c;
};
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
struct Xf {
  V2 p;
  Rot q;
};

// src/b2_math.rs:611-680
B2G_HD V2 mul(Rot q, V2 v) { return v2(q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y); }
B2G_HD V2 mulT(Rot q, V2 v) { return v2(q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y); }
B2G_HD Rot mulT(Rot q, Rot r) { Rot o; o.s = q.c * r.s - q.s * r.c; o.c = q.c * r.c + q.s * r.s; return o; }
B2G_HD V2 mul(Xf t, V2 v) {
  float x = (t.q.c * v.x - t.q.s * v.y) + t.p.x;
  float y = (t.q.s * v.x + t.q.c * v.y) + t.p.y;
  return v2(x, y);
}
B2G_HD V2 mulT(Xf t, V2 v) {
  float px = v.x - t.p.x, py = v.y - t.p.y;
  return v2(t.q.c * px + t.q.s * py, -t.q.s * px + t.q.c * py);
}
B2G_HD Xf mulT(Xf a, Xf b) {
  Xf c;
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
