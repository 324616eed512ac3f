// CPU oracle for the box2d-rs B2world::step hot path -- TEST INFRASTRUCTURE ONLY.
//
// This is a plain sequential C++ restatement of the reference algorithm, written so that the
// GPU implementation can be checked against it. It is never linked into, loaded by, or
// called from the product library; only the tests, the smoke check and the CPU-baseline leg
// of the benchmark use it, and only as the checker.
//
// Part 1: math, shapes and narrowphase collision. Every function follows the reference
// operation by operation (no FMA contraction: build with -ffp-contract=off) so that float
// results round identically. Citations are to the reference crate (box2d-rs 0.0.4).
//
// `real` is float by default; building with -DB2O_REAL=double gives the f64 instantiation
// used for finite-difference checks.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

#ifndef B2O_REAL
#define B2O_REAL float
#endif

namespace b2o {

typedef B2O_REAL real;

// ---- constants: src/b2_common.rs:17-91, src/b2_settings.rs:7-11 ---------------------------
static const real kMaxFloat = std::numeric_limits<real>::max();
static const real kEpsilon = std::numeric_limits<real>::epsilon();
static const real kPi = real(3.14159265358979323846);
static const int kMaxManifoldPoints = 2;
static const int kMaxPolygonVertices = 8;
static const real kAabbExtension = real(0.1);
static const real kAabbMultiplier = real(4.0);
static const real kLinearSlop = real(0.005);
static const real kPolygonRadius = real(2.0) * kLinearSlop;
static const real kMaxLinearCorrection = real(0.2);
static const real kMaxTranslation = real(2.0);
static const real kMaxTranslationSquared = kMaxTranslation * kMaxTranslation;
static const real kMaxRotation = real(0.5) * kPi;
static const real kMaxRotationSquared = kMaxRotation * kMaxRotation;
static const real kBaumgarte = real(0.2);
static const real kToiBaumgarte = real(0.75);  // src/b2_common.rs:80
static const int kMaxSubSteps = 8;             // src/b2_common.rs:51
static const real kTimeToSleep = real(0.5);
static const real kLinearSleepTolerance = real(0.01);
static const real kAngularSleepTolerance = real(2.0) / real(180.0) * kPi;

// ---- math: src/b2_math.rs ----------------------------------------------------------------
struct Vec2 {
  real x, y;
  Vec2() : x(0), y(0) {}
  Vec2(real x_, real y_) : x(x_), y(y_) {}
  void set(real x_, real y_) { x = x_; y = y_; }
  void set_zero() { x = 0; y = 0; }
  // src/b2_math.rs:71-73
  real length() const { return std::sqrt(x * x + y * y); }
  real length_squared() const { return x * x + y * y; }
  // src/b2_math.rs:82-92
  real normalize() {
    real len = length();
    if (len < kEpsilon) return 0;
    real inv = real(1.0) / len;
    x *= inv;
    y *= inv;
    return len;
  }
  Vec2 operator-() const { return Vec2(-x, -y); }
  void operator+=(Vec2 o) { x += o.x; y += o.y; }
  void operator-=(Vec2 o) { x -= o.x; y -= o.y; }
  void operator*=(real s) { x *= s; y *= s; }
};
inline Vec2 operator+(Vec2 a, Vec2 b) { return Vec2(a.x + b.x, a.y + b.y); }
inline Vec2 operator-(Vec2 a, Vec2 b) { return Vec2(a.x - b.x, a.y - b.y); }
inline Vec2 operator*(real s, Vec2 a) { return Vec2(s * a.x, s * a.y); }
inline bool operator==(Vec2 a, Vec2 b) { return a.x == b.x && a.y == b.y; }

// src/b2_math.rs:470-491
inline real dot(Vec2 a, Vec2 b) { return a.x * b.x + a.y * b.y; }
inline real cross(Vec2 a, Vec2 b) { return a.x * b.y - a.y * b.x; }
inline Vec2 cross(Vec2 a, real s) { return Vec2(s * a.y, -s * a.x); }
inline Vec2 cross(real s, Vec2 a) { return Vec2(-s * a.y, s * a.x); }
// src/b2_math.rs:533-541
inline real distance_squared(Vec2 a, Vec2 b) { Vec2 c = a - b; return dot(c, c); }
// src/b2_math.rs:704-724 (b2_min/b2_max are `a < b ? a : b` / `a > b ? a : b`)
template <class T> inline T bmin(T a, T b) { return a < b ? a : b; }
template <class T> inline T bmax(T a, T b) { return a > b ? a : b; }
template <class T> inline T bclamp(T a, T lo, T hi) { return bmax(lo, bmin(a, hi)); }
inline Vec2 vmin(Vec2 a, Vec2 b) { return Vec2(bmin(a.x, b.x), bmin(a.y, b.y)); }
inline Vec2 vmax(Vec2 a, Vec2 b) { return Vec2(bmax(a.x, b.x), bmax(a.y, b.y)); }
inline real babs(real v) { return std::fabs(v); }

// src/b2_math.rs:216-296
struct Mat22 {
  Vec2 ex, ey;
  void set_zero() { ex.set_zero(); ey.set_zero(); }
  Mat22 get_inverse() const {
    real a = ex.x, b = ey.x, c = ex.y, d = ey.y;
    real det = a * d - b * c;
    if (det != 0) det = real(1.0) / det;
    Mat22 r;
    r.ex = Vec2(det * d, -det * c);
    r.ey = Vec2(-det * b, det * a);
    return r;
  }
};
// src/b2_math.rs:493-495
inline Vec2 mul(const Mat22& a, Vec2 v) { return Vec2(a.ex.x * v.x + a.ey.x * v.y, a.ex.y * v.x + a.ey.y * v.y); }

// src/b2_math.rs:355-400
struct Rot {
  real s, c;
  Rot() : s(0), c(1) {}
  explicit Rot(real angle) { set(angle); }
  void set(real angle) { s = std::sin(angle); c = std::cos(angle); }
  real get_angle() const { return std::atan2(s, c); }
};
// src/b2_math.rs:404-432
struct Transform {
  Vec2 p;
  Rot q;
  Transform() {}
  Transform(Vec2 p_, Rot q_) : p(p_), q(q_) {}
};
// src/b2_math.rs:611-680
inline Rot mul(Rot q, Rot r) { Rot o; o.s = q.s * r.c + q.c * r.s; o.c = q.c * r.c - q.s * r.s; return o; }
inline Rot mulT(Rot q, Rot r) { Rot o; o.s = q.c * r.s - q.s * r.c; o.c = q.c * r.c + q.s * r.s; return o; }
inline Vec2 mul(Rot q, Vec2 v) { return Vec2(q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y); }
inline Vec2 mulT(Rot q, Vec2 v) { return Vec2(q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y); }
inline Vec2 mul(const Transform& t, Vec2 v) {
  real x = (t.q.c * v.x - t.q.s * v.y) + t.p.x;
  real y = (t.q.s * v.x + t.q.c * v.y) + t.p.y;
  return Vec2(x, y);
}
inline Vec2 mulT(const Transform& t, Vec2 v) {
  real px = v.x - t.p.x, py = v.y - t.p.y;
  return Vec2(t.q.c * px + t.q.s * py, -t.q.s * px + t.q.c * py);
}
inline Transform mulT(const Transform& a, const Transform& b) {
  Transform c;
  c.q = mulT(a.q, b.q);
  c.p = mulT(a.q, b.p - a.p);
  return c;
}

// src/b2_math.rs:437-447, 764-787
struct Sweep {
  Vec2 local_center, c0, c;
  real a0 = 0, a = 0, alpha0 = 0;
  void get_transform(Transform* xf, real beta) const {
    xf->p = (real(1.0) - beta) * c0 + beta * c;
    real angle = (real(1.0) - beta) * a0 + beta * a;
    xf->q.set(angle);
    xf->p -= mul(xf->q, local_center);
  }
};

// ---- AABB: src/b2_collision.rs:201-256, 355-368 ------------------------------------------
struct AABB {
  Vec2 lower, upper;
  real perimeter() const {
    real wx = upper.x - lower.x, wy = upper.y - lower.y;
    return real(2.0) * (wx + wy);
  }
  Vec2 center() const { return real(0.5) * (lower + upper); }
  void combine(const AABB& a, const AABB& b) { lower = vmin(a.lower, b.lower); upper = vmax(a.upper, b.upper); }
  bool contains(const AABB& o) const {
    bool r = true;
    r = r && lower.x <= o.lower.x;
    r = r && lower.y <= o.lower.y;
    r = r && o.upper.x <= upper.x;
    r = r && o.upper.y <= upper.y;
    return r;
  }
};
inline bool test_overlap(const AABB& a, const AABB& b) {
  Vec2 d1 = b.lower - a.upper, d2 = a.lower - b.upper;
  if (d1.x > 0 || d1.y > 0) return false;
  if (d2.x > 0 || d2.y > 0) return false;
  return true;
}

// ---- shapes: src/b2_shape.rs:25-31, src/private/collision/b2_*_shape.rs -------------------
enum ShapeType { kCircle = 0, kEdge = 1, kPolygon = 2, kChain = 3 };

struct MassData { real mass = 0; Vec2 center; real I = 0; };

struct Shape {
  int type = kCircle;
  real radius = 0;
  // circle
  Vec2 p;
  // polygon
  int count = 0;
  Vec2 vertices[kMaxPolygonVertices];
  Vec2 normals[kMaxPolygonVertices];
  Vec2 centroid;
  // edge
  Vec2 v0, v1, v2, v3;
  bool one_sided = false;
};

// src/private/collision/b2_polygon_shape.rs:12-23
inline void set_as_box(Shape* s, real hx, real hy) {
  s->type = kPolygon;
  s->radius = kPolygonRadius;
  s->count = 4;
  s->vertices[0].set(-hx, -hy);
  s->vertices[1].set(hx, -hy);
  s->vertices[2].set(hx, hy);
  s->vertices[3].set(-hx, hy);
  s->normals[0].set(0, -1);
  s->normals[1].set(1, 0);
  s->normals[2].set(0, 1);
  s->normals[3].set(-1, 0);
  s->centroid.set_zero();
}

// src/private/collision/b2_polygon_shape.rs:25-52
inline void set_as_box_angle(Shape* s, real hx, real hy, Vec2 center, real angle) {
  set_as_box(s, hx, hy);
  s->centroid = center;
  Transform xf(center, Rot(angle));
  for (int i = 0; i < s->count; ++i) {
    s->vertices[i] = mul(xf, s->vertices[i]);
    s->normals[i] = mul(xf.q, s->normals[i]);
  }
}

// src/private/collision/b2_polygon_shape.rs:59-94
inline Vec2 compute_centroid(const Vec2* vs, int count) {
  Vec2 c;
  real area = 0;
  Vec2 s = vs[0];
  const real inv3 = real(1.0) / real(3.0);
  for (int i = 0; i < count; ++i) {
    Vec2 p1 = vs[0] - s;
    Vec2 p2 = vs[i] - s;
    Vec2 p3 = i + 1 < count ? vs[i + 1] - s : vs[0] - s;
    Vec2 e1 = p2 - p1, e2 = p3 - p1;
    real d = cross(e1, e2);
    real tri = real(0.5) * d;
    area += tri;
    c += (tri * inv3) * ((p1 + p2) + p3);
  }
  c = (real(1.0) / area) * c + s;
  return c;
}

// Gift-wrapping hull: src/private/collision/b2_polygon_shape.rs:96-211.
// Returns false when the point set is degenerate (the reference then asserts).
inline bool polygon_set(Shape* s, const Vec2* verts, int count) {
  s->type = kPolygon;
  s->radius = kPolygonRadius;
  if (count < 3) { set_as_box(s, 1, 1); return false; }
  int n = bmin(count, kMaxPolygonVertices);
  Vec2 ps[kMaxPolygonVertices];
  int temp_count = 0;
  for (int i = 0; i < n; ++i) {
    Vec2 v = verts[i];
    bool unique = true;
    for (int j = 0; j < temp_count; ++j) {
      if (distance_squared(v, ps[j]) < (real(0.5) * kLinearSlop) * (real(0.5) * kLinearSlop)) { unique = false; break; }
    }
    if (unique) ps[temp_count++] = v;
  }
  n = temp_count;
  if (n < 3) { set_as_box(s, 1, 1); return false; }
  int i0 = 0;
  real x0 = ps[0].x;
  for (int i = 1; i < n; ++i) {
    real x = ps[i].x;
    if (x > x0 || (x == x0 && ps[i].y < ps[i0].y)) { i0 = i; x0 = x; }
  }
  int hull[kMaxPolygonVertices];
  int m = 0, ih = i0;
  for (;;) {
    if (m >= kMaxPolygonVertices) { set_as_box(s, 1, 1); return false; }
    hull[m] = ih;
    int ie = 0;
    for (int j = 1; j < n; ++j) {
      if (ie == ih) { ie = j; continue; }
      Vec2 r = ps[ie] - ps[hull[m]];
      Vec2 v = ps[j] - ps[hull[m]];
      real c = cross(r, v);
      if (c < 0) ie = j;
      if (c == 0 && v.length_squared() > r.length_squared()) ie = j;
    }
    ++m;
    ih = ie;
    if (ie == i0) break;
  }
  if (m < 3) { set_as_box(s, 1, 1); return false; }
  s->count = m;
  for (int i = 0; i < m; ++i) s->vertices[i] = ps[hull[i]];
  for (int i = 0; i < m; ++i) {
    int i2 = i + 1 < m ? i + 1 : 0;
    Vec2 edge = s->vertices[i2] - s->vertices[i];
    s->normals[i] = cross(edge, real(1.0));
    s->normals[i].normalize();
  }
  s->centroid = compute_centroid(s->vertices, m);
  return true;
}

inline void make_circle(Shape* s, Vec2 p, real r) { s->type = kCircle; s->p = p; s->radius = r; }
// src/private/collision/b2_edge_shape.rs:15-19 and the one-sided setter just above it
inline void make_edge_two_sided(Shape* s, Vec2 a, Vec2 b) {
  s->type = kEdge; s->radius = kPolygonRadius; s->v1 = a; s->v2 = b; s->one_sided = false;
}
inline void make_edge_one_sided(Shape* s, Vec2 v0, Vec2 v1, Vec2 v2, Vec2 v3) {
  s->type = kEdge; s->radius = kPolygonRadius; s->v0 = v0; s->v1 = v1; s->v2 = v2; s->v3 = v3; s->one_sided = true;
}

// compute_aabb: polygon src/private/collision/b2_polygon_shape.rs:292-312,
// circle src/private/collision/b2_circle_shape.rs:64-77, edge src/private/collision/b2_edge_shape.rs:104-124
inline void compute_aabb(const Shape& s, AABB* aabb, const Transform& xf) {
  if (s.type == kCircle) {
    Vec2 p = xf.p + mul(xf.q, s.p);
    aabb->lower.set(p.x - s.radius, p.y - s.radius);
    aabb->upper.set(p.x + s.radius, p.y + s.radius);
  } else if (s.type == kPolygon) {
    Vec2 lower = mul(xf, s.vertices[0]);
    Vec2 upper = lower;
    for (int i = 1; i < s.count; ++i) {
      Vec2 v = mul(xf, s.vertices[i]);
      lower = vmin(lower, v);
      upper = vmax(upper, v);
    }
    Vec2 r(s.radius, s.radius);
    aabb->lower = lower - r;
    aabb->upper = upper + r;
  } else {
    Vec2 v1 = mul(xf, s.v1), v2 = mul(xf, s.v2);
    Vec2 lower = vmin(v1, v2), upper = vmax(v1, v2);
    Vec2 r(s.radius, s.radius);
    aabb->lower = lower - r;
    aabb->upper = upper + r;
  }
}

// compute_mass: polygon src/private/collision/b2_polygon_shape.rs:314-390,
// circle src/private/collision/b2_circle_shape.rs:79-86, edge src/private/collision/b2_edge_shape.rs:126-132
inline void compute_mass(const Shape& s, MassData* md, real density) {
  if (s.type == kCircle) {
    md->mass = density * kPi * s.radius * s.radius;
    md->center = s.p;
    md->I = md->mass * (real(0.5) * s.radius * s.radius + dot(s.p, s.p));
  } else if (s.type == kPolygon) {
    Vec2 center;
    real area = 0, inert = 0;
    Vec2 sref = s.vertices[0];
    const real inv3 = real(1.0) / real(3.0);
    for (int i = 0; i < s.count; ++i) {
      Vec2 e1 = s.vertices[i] - sref;
      Vec2 e2 = i + 1 < s.count ? s.vertices[i + 1] - sref : s.vertices[0] - sref;
      real d = cross(e1, e2);
      real tri = real(0.5) * d;
      area += tri;
      center += (tri * inv3) * (e1 + e2);
      real ex1 = e1.x, ey1 = e1.y, ex2 = e2.x, ey2 = e2.y;
      real intx2 = ex1 * ex1 + ex2 * ex1 + ex2 * ex2;
      real inty2 = ey1 * ey1 + ey2 * ey1 + ey2 * ey2;
      inert += ((real(0.25) * inv3) * d) * (intx2 + inty2);
    }
    md->mass = density * area;
    center *= real(1.0) / area;
    md->center = center + sref;
    md->I = density * inert;
    md->I += md->mass * (dot(md->center, md->center) - dot(center, center));
  } else {
    md->mass = 0;
    md->center = real(0.5) * (s.v1 + s.v2);
    md->I = 0;
  }
}

// ---- manifold types: src/b2_collision.rs:23-114 ------------------------------------------
enum ManifoldType { kCircles = 0, kFaceA = 1, kFaceB = 2 };
enum FeatureType { kVertex = 0, kFace = 1 };

struct ContactFeature {
  uint8_t index_a = 0, index_b = 0, type_a = 0, type_b = 0;
  uint32_t key() const { return uint32_t(index_a) | uint32_t(index_b) << 8 | uint32_t(type_a) << 16 | uint32_t(type_b) << 24; }
  bool operator==(const ContactFeature& o) const {
    return index_a == o.index_a && index_b == o.index_b && type_a == o.type_a && type_b == o.type_b;
  }
};

struct ManifoldPoint {
  Vec2 local_point;
  real normal_impulse = 0, tangent_impulse = 0;
  ContactFeature id;
};

struct Manifold {
  ManifoldPoint points[kMaxManifoldPoints];
  Vec2 local_normal, local_point;
  int type = kCircles;
  int point_count = 0;
};

struct WorldManifold {
  Vec2 normal;
  Vec2 points[kMaxManifoldPoints];
  real separations[kMaxManifoldPoints] = {0, 0};
};

// src/private/collision/b2_collision.rs:7-68
inline void world_manifold_initialize(WorldManifold* wm, const Manifold& m, const Transform& xfA, real rA,
                                      const Transform& xfB, real rB) {
  if (m.point_count == 0) return;
  if (m.type == kCircles) {
    wm->normal.set(1, 0);
    Vec2 pA = mul(xfA, m.local_point);
    Vec2 pB = mul(xfB, m.points[0].local_point);
    if (distance_squared(pA, pB) > kEpsilon * kEpsilon) {
      wm->normal = pB - pA;
      wm->normal.normalize();
    }
    Vec2 cA = pA + rA * wm->normal;
    Vec2 cB = pB - rB * wm->normal;
    wm->points[0] = real(0.5) * (cA + cB);
    wm->separations[0] = dot(cB - cA, wm->normal);
  } else if (m.type == kFaceA) {
    wm->normal = mul(xfA.q, m.local_normal);
    Vec2 plane = mul(xfA, m.local_point);
    for (int i = 0; i < m.point_count; ++i) {
      Vec2 clip = mul(xfB, m.points[i].local_point);
      Vec2 cA = clip + (rA - dot(clip - plane, wm->normal)) * wm->normal;
      Vec2 cB = clip - rB * wm->normal;
      wm->points[i] = real(0.5) * (cA + cB);
      wm->separations[i] = dot(cB - cA, wm->normal);
    }
  } else {
    wm->normal = mul(xfB.q, m.local_normal);
    Vec2 plane = mul(xfB, m.local_point);
    for (int i = 0; i < m.point_count; ++i) {
      Vec2 clip = mul(xfA, m.points[i].local_point);
      Vec2 cB = clip + (rB - dot(clip - plane, wm->normal)) * wm->normal;
      Vec2 cA = clip - rA * wm->normal;
      wm->points[i] = real(0.5) * (cA + cB);
      wm->separations[i] = dot(cA - cB, wm->normal);
    }
    wm->normal = -wm->normal;
  }
}

struct ClipVertex { Vec2 v; ContactFeature id; };

// Sutherland-Hodgman clipping: src/private/collision/b2_collision.rs:171-214
inline int clip_segment_to_line(ClipVertex vOut[2], const ClipVertex vIn[2], Vec2 normal, real offset, int vertexIndexA) {
  int count = 0;
  real d0 = dot(normal, vIn[0].v) - offset;
  real d1 = dot(normal, vIn[1].v) - offset;
  if (d0 <= 0) vOut[count++] = vIn[0];
  if (d1 <= 0) vOut[count++] = vIn[1];
  if (d0 * d1 < 0) {
    real interp = d0 / (d0 - d1);
    vOut[count].v = vIn[0].v + interp * (vIn[1].v - vIn[0].v);
    vOut[count].id.index_a = uint8_t(vertexIndexA);
    vOut[count].id.index_b = vIn[0].id.index_b;
    vOut[count].id.type_a = kVertex;
    vOut[count].id.type_b = kFace;
    ++count;
  }
  return count;
}

// src/private/collision/b2_collide_circle.rs:7-34
inline void collide_circles(Manifold* m, const Shape& a, const Transform& xfA, const Shape& b, const Transform& xfB) {
  m->point_count = 0;
  Vec2 pA = mul(xfA, a.p), pB = mul(xfB, b.p);
  Vec2 d = pB - pA;
  real dist_sqr = dot(d, d);
  real radius = a.radius + b.radius;
  if (dist_sqr > radius * radius) return;
  m->type = kCircles;
  m->local_point = a.p;
  m->local_normal.set_zero();
  m->point_count = 1;
  m->points[0].local_point = b.p;
  m->points[0].id = ContactFeature();
}

// src/private/collision/b2_collide_circle.rs:36-133
inline void collide_polygon_circle(Manifold* m, const Shape& poly, const Transform& xfA, const Shape& circ,
                                   const Transform& xfB) {
  m->point_count = 0;
  Vec2 c = mul(xfB, circ.p);
  Vec2 cLocal = mulT(xfA, c);
  int normal_index = 0;
  real separation = -kMaxFloat;
  real radius = poly.radius + circ.radius;
  int vc = poly.count;
  const Vec2* vs = poly.vertices;
  const Vec2* ns = poly.normals;
  for (int i = 0; i < vc; ++i) {
    real s = dot(ns[i], cLocal - vs[i]);
    if (s > radius) return;
    if (s > separation) { separation = s; normal_index = i; }
  }
  int i1 = normal_index, i2 = i1 + 1 < vc ? i1 + 1 : 0;
  Vec2 v1 = vs[i1], v2 = vs[i2];
  if (separation < kEpsilon) {
    m->point_count = 1;
    m->type = kFaceA;
    m->local_normal = ns[normal_index];
    m->local_point = real(0.5) * (v1 + v2);
    m->points[0].local_point = circ.p;
    m->points[0].id = ContactFeature();
    return;
  }
  real u1 = dot(cLocal - v1, v2 - v1);
  real u2 = dot(cLocal - v2, v1 - v2);
  if (u1 <= 0) {
    if (distance_squared(cLocal, v1) > radius * radius) return;
    m->point_count = 1;
    m->type = kFaceA;
    m->local_normal = cLocal - v1;
    m->local_normal.normalize();
    m->local_point = v1;
    m->points[0].local_point = circ.p;
    m->points[0].id = ContactFeature();
  } else if (u2 <= 0) {
    if (distance_squared(cLocal, v2) > radius * radius) return;
    m->point_count = 1;
    m->type = kFaceA;
    m->local_normal = cLocal - v2;
    m->local_normal.normalize();
    m->local_point = v2;
    m->points[0].local_point = circ.p;
    m->points[0].id = ContactFeature();
  } else {
    Vec2 fc = real(0.5) * (v1 + v2);
    real s = dot(cLocal - fc, ns[i1]);
    if (s > radius) return;
    m->point_count = 1;
    m->type = kFaceA;
    m->local_normal = ns[i1];
    m->local_point = fc;
    m->points[0].local_point = circ.p;
    m->points[0].id = ContactFeature();
  }
}

// src/private/collision/b2_collide_polygon.rs:8-46
inline real find_max_separation(int* edgeIndex, const Shape& p1, const Transform& xf1, const Shape& p2, const Transform& xf2) {
  int c1 = p1.count, c2 = p2.count;
  Transform xf = mulT(xf2, xf1);
  int best = 0;
  real maxSep = -kMaxFloat;
  for (int i = 0; i < c1; ++i) {
    Vec2 n = mul(xf.q, p1.normals[i]);
    Vec2 v1 = mul(xf, p1.vertices[i]);
    real si = kMaxFloat;
    for (int j = 0; j < c2; ++j) {
      real sij = dot(n, p2.vertices[j] - v1);
      if (sij < si) si = sij;
    }
    if (si > maxSep) { maxSep = si; best = i; }
  }
  *edgeIndex = best;
  return maxSep;
}

// src/private/collision/b2_collide_polygon.rs:48-93
inline void find_incident_edge(ClipVertex c[2], const Shape& p1, const Transform& xf1, int edge1, const Shape& p2,
                               const Transform& xf2) {
  int count2 = p2.count;
  Vec2 normal1 = mulT(xf2.q, mul(xf1.q, p1.normals[edge1]));
  int index = 0;
  real minDot = kMaxFloat;
  for (int i = 0; i < count2; ++i) {
    real d = dot(normal1, p2.normals[i]);
    if (d < minDot) { minDot = d; index = i; }
  }
  int i1 = index, i2 = i1 + 1 < count2 ? i1 + 1 : 0;
  c[0].v = mul(xf2, p2.vertices[i1]);
  c[0].id.index_a = uint8_t(edge1); c[0].id.index_b = uint8_t(i1); c[0].id.type_a = kFace; c[0].id.type_b = kVertex;
  c[1].v = mul(xf2, p2.vertices[i2]);
  c[1].id.index_a = uint8_t(edge1); c[1].id.index_b = uint8_t(i2); c[1].id.type_a = kFace; c[1].id.type_b = kVertex;
}

// src/private/collision/b2_collide_polygon.rs:102-225
inline void collide_polygons(Manifold* m, const Shape& pA, const Transform& xfA, const Shape& pB, const Transform& xfB) {
  m->point_count = 0;
  real totalRadius = pA.radius + pB.radius;
  int edgeA = 0;
  real sepA = find_max_separation(&edgeA, pA, xfA, pB, xfB);
  if (sepA > totalRadius) return;
  int edgeB = 0;
  real sepB = find_max_separation(&edgeB, pB, xfB, pA, xfA);
  if (sepB > totalRadius) return;
  const Shape *poly1, *poly2;
  Transform xf1, xf2;
  int edge1;
  uint8_t flip;
  const real kTol = real(0.1) * kLinearSlop;
  if (sepB > sepA + kTol) {
    poly1 = &pB; poly2 = &pA; xf1 = xfB; xf2 = xfA; edge1 = edgeB; m->type = kFaceB; flip = 1;
  } else {
    poly1 = &pA; poly2 = &pB; xf1 = xfA; xf2 = xfB; edge1 = edgeA; m->type = kFaceA; flip = 0;
  }
  ClipVertex incident[2];
  find_incident_edge(incident, *poly1, xf1, edge1, *poly2, xf2);
  int count1 = poly1->count;
  int iv1 = edge1, iv2 = edge1 + 1 < count1 ? edge1 + 1 : 0;
  Vec2 v11 = poly1->vertices[iv1], v12 = poly1->vertices[iv2];
  Vec2 localTangent = v12 - v11;
  localTangent.normalize();
  Vec2 localNormal = cross(localTangent, real(1.0));
  Vec2 planePoint = real(0.5) * (v11 + v12);
  Vec2 tangent = mul(xf1.q, localTangent);
  Vec2 normal = cross(tangent, real(1.0));
  v11 = mul(xf1, v11);
  v12 = mul(xf1, v12);
  real frontOffset = dot(normal, v11);
  real sideOffset1 = -dot(tangent, v11) + totalRadius;
  real sideOffset2 = dot(tangent, v12) + totalRadius;
  ClipVertex cp1[2], cp2[2];
  int np = clip_segment_to_line(cp1, incident, -tangent, sideOffset1, iv1);
  if (np < 2) return;
  np = clip_segment_to_line(cp2, cp1, tangent, sideOffset2, iv2);
  if (np < 2) return;
  m->local_normal = localNormal;
  m->local_point = planePoint;
  int pc = 0;
  for (int i = 0; i < kMaxManifoldPoints; ++i) {
    real separation = dot(normal, cp2[i].v) - frontOffset;
    if (separation <= totalRadius) {
      ManifoldPoint* cp = &m->points[pc];
      cp->local_point = mulT(xf2, cp2[i].v);
      cp->id = cp2[i].id;
      if (flip) {
        ContactFeature cf = cp->id;
        cp->id.index_a = cf.index_b; cp->id.index_b = cf.index_a;
        cp->id.type_a = cf.type_b; cp->id.type_b = cf.type_a;
      }
      ++pc;
    }
  }
  m->point_count = pc;
}

// src/private/collision/b2_collide_edge.rs:11-123
inline void collide_edge_circle(Manifold* m, const Shape& edge, const Transform& xfA, const Shape& circ, const Transform& xfB) {
  m->point_count = 0;
  Vec2 Q = mulT(xfA, mul(xfB, circ.p));
  Vec2 A = edge.v1, B = edge.v2;
  Vec2 e = B - A;
  Vec2 n(e.y, -e.x);
  real offset = dot(n, Q - A);
  bool oneSided = edge.one_sided;
  if (oneSided && offset < 0) return;
  real u = dot(e, B - Q);
  real v = dot(e, Q - A);
  real radius = edge.radius + circ.radius;
  ContactFeature cf;
  cf.index_b = 0;
  cf.type_b = kVertex;
  if (v <= 0) {
    Vec2 P = A;
    Vec2 d = Q - P;
    real dd = dot(d, d);
    if (dd > radius * radius) return;
    if (edge.one_sided) {
      Vec2 A1 = edge.v0, B1 = A;
      Vec2 e1 = B1 - A1;
      real u1 = dot(e1, B1 - Q);
      if (u1 > 0) return;
    }
    cf.index_a = 0; cf.type_a = kVertex;
    m->point_count = 1; m->type = kCircles; m->local_normal.set_zero(); m->local_point = P;
    m->points[0].id = cf; m->points[0].local_point = circ.p;
    return;
  }
  if (u <= 0) {
    Vec2 P = B;
    Vec2 d = Q - P;
    real dd = dot(d, d);
    if (dd > radius * radius) return;
    if (edge.one_sided) {
      Vec2 B2 = edge.v3, A2 = B;
      Vec2 e2 = B2 - A2;
      real v2 = dot(e2, Q - A2);
      if (v2 > 0) return;
    }
    cf.index_a = 1; cf.type_a = kVertex;
    m->point_count = 1; m->type = kCircles; m->local_normal.set_zero(); m->local_point = P;
    m->points[0].id = cf; m->points[0].local_point = circ.p;
    return;
  }
  real den = dot(e, e);
  Vec2 P = (real(1.0) / den) * (u * A + v * B);
  Vec2 d = Q - P;
  real dd = dot(d, d);
  if (dd > radius * radius) return;
  if (offset < 0) n.set(-n.x, -n.y);
  n.normalize();
  cf.index_a = 0; cf.type_a = kFace;
  m->point_count = 1; m->type = kFaceA; m->local_normal = n; m->local_point = A;
  m->points[0].id = cf; m->points[0].local_point = circ.p;
}

// src/private/collision/b2_collide_edge.rs:125-228
enum EPAxisType { kUnknown = 0, kEdgeA = 1, kEdgeB = 2 };
struct EPAxis { Vec2 normal; int type = kUnknown; int index = -1; real separation = 0; };
struct TempPolygon { Vec2 vertices[kMaxPolygonVertices]; Vec2 normals[kMaxPolygonVertices]; int count = 0; };
struct ReferenceFace {
  int i1 = 0, i2 = 0;
  Vec2 v1, v2, normal, sideNormal1, sideNormal2;
  real sideOffset1 = 0, sideOffset2 = 0;
};

inline EPAxis compute_edge_separation(const TempPolygon& pb, Vec2 v1, Vec2 normal1) {
  EPAxis axis;
  axis.type = kEdgeA; axis.index = -1; axis.separation = -kMaxFloat; axis.normal = Vec2();
  Vec2 axes[2] = {normal1, -normal1};
  for (int j = 0; j < 2; ++j) {
    real sj = kMaxFloat;
    for (int i = 0; i < pb.count; ++i) {
      real si = dot(axes[j], pb.vertices[i] - v1);
      if (si < sj) sj = si;
    }
    if (sj > axis.separation) { axis.index = j; axis.separation = sj; axis.normal = axes[j]; }
  }
  return axis;
}

inline EPAxis compute_polygon_separation(const TempPolygon& pb, Vec2 v1, Vec2 v2) {
  EPAxis axis;
  axis.type = kUnknown; axis.index = -1; axis.separation = -kMaxFloat; axis.normal = Vec2();
  for (int i = 0; i < pb.count; ++i) {
    Vec2 n = -pb.normals[i];
    real s1 = dot(n, pb.vertices[i] - v1);
    real s2 = dot(n, pb.vertices[i] - v2);
    real s = bmin(s1, s2);
    if (s > axis.separation) { axis.type = kEdgeB; axis.index = i; axis.separation = s; axis.normal = n; }
  }
  return axis;
}

// src/private/collision/b2_collide_edge.rs:230-475
inline void collide_edge_polygon(Manifold* m, const Shape& edge, const Transform& xfA, const Shape& poly, const Transform& xfB) {
  m->point_count = 0;
  Transform xf = mulT(xfA, xfB);
  Vec2 centroidB = mul(xf, poly.centroid);
  Vec2 v1 = edge.v1, v2 = edge.v2;
  Vec2 edge1 = v2 - v1;
  edge1.normalize();
  Vec2 normal1(edge1.y, -edge1.x);
  real offset1 = dot(normal1, centroidB - v1);
  bool oneSided = edge.one_sided;
  if (oneSided && offset1 < 0) return;
  TempPolygon tp;
  tp.count = poly.count;
  for (int i = 0; i < poly.count; ++i) {
    tp.vertices[i] = mul(xf, poly.vertices[i]);
    tp.normals[i] = mul(xf.q, poly.normals[i]);
  }
  real radius = poly.radius + edge.radius;
  EPAxis edgeAxis = compute_edge_separation(tp, v1, normal1);
  if (edgeAxis.separation > radius) return;
  EPAxis polygonAxis = compute_polygon_separation(tp, v1, v2);
  if (polygonAxis.separation > radius) return;
  const real kRelativeTol = real(0.98), kAbsoluteTol = real(0.001);
  EPAxis primary;
  if (polygonAxis.separation - radius > kRelativeTol * (edgeAxis.separation - radius) + kAbsoluteTol) primary = polygonAxis;
  else primary = edgeAxis;
  if (oneSided) {
    Vec2 edge0 = v1 - edge.v0;
    edge0.normalize();
    Vec2 normal0(edge0.y, -edge0.x);
    bool convex1 = cross(edge0, edge1) >= 0;
    Vec2 edge2 = edge.v3 - v2;
    edge2.normalize();
    Vec2 normal2(edge2.y, -edge2.x);
    bool convex2 = cross(edge1, edge2) >= 0;
    const real sinTol = real(0.1);
    bool side1 = dot(primary.normal, edge1) <= 0;
    if (side1) {
      if (convex1) {
        if (cross(primary.normal, normal0) > sinTol) return;
      } else {
        primary = edgeAxis;
      }
    } else {
      if (convex2) {
        if (cross(normal2, primary.normal) > sinTol) return;
      } else {
        primary = edgeAxis;
      }
    }
  }
  ClipVertex clip[2];
  ReferenceFace rf;
  if (primary.type == kEdgeA) {
    m->type = kFaceA;
    int bestIndex = 0;
    real bestValue = dot(primary.normal, tp.normals[0]);
    for (int i = 1; i < tp.count; ++i) {
      real value = dot(primary.normal, tp.normals[i]);
      if (value < bestValue) { bestValue = value; bestIndex = i; }
    }
    int i1 = bestIndex, i2 = i1 + 1 < tp.count ? i1 + 1 : 0;
    clip[0].v = tp.vertices[i1];
    clip[0].id.index_a = 0; clip[0].id.index_b = uint8_t(i1); clip[0].id.type_a = kFace; clip[0].id.type_b = kVertex;
    clip[1].v = tp.vertices[i2];
    clip[1].id.index_a = 0; clip[1].id.index_b = uint8_t(i2); clip[1].id.type_a = kFace; clip[1].id.type_b = kVertex;
    rf.i1 = 0; rf.i2 = 1; rf.v1 = v1; rf.v2 = v2; rf.normal = primary.normal;
    rf.sideNormal1 = -edge1; rf.sideNormal2 = edge1;
  } else {
    m->type = kFaceB;
    clip[0].v = v2;
    clip[0].id.index_a = 1; clip[0].id.index_b = uint8_t(primary.index); clip[0].id.type_a = kVertex; clip[0].id.type_b = kFace;
    clip[1].v = v1;
    clip[1].id.index_a = 0; clip[1].id.index_b = uint8_t(primary.index); clip[1].id.type_a = kVertex; clip[1].id.type_b = kFace;
    rf.i1 = primary.index;
    rf.i2 = rf.i1 + 1 < tp.count ? rf.i1 + 1 : 0;
    rf.v1 = tp.vertices[rf.i1]; rf.v2 = tp.vertices[rf.i2];
    rf.normal = tp.normals[rf.i1];
    rf.sideNormal1.set(rf.normal.y, -rf.normal.x);
    rf.sideNormal2 = -rf.sideNormal1;
  }
  rf.sideOffset1 = dot(rf.sideNormal1, rf.v1);
  rf.sideOffset2 = dot(rf.sideNormal2, rf.v2);
  ClipVertex c1[2], c2[2];
  int np = clip_segment_to_line(c1, clip, rf.sideNormal1, rf.sideOffset1, rf.i1);
  if (np < kMaxManifoldPoints) return;
  np = clip_segment_to_line(c2, c1, rf.sideNormal2, rf.sideOffset2, rf.i2);
  if (np < kMaxManifoldPoints) return;
  if (primary.type == kEdgeA) {
    m->local_normal = rf.normal;
    m->local_point = rf.v1;
  } else {
    m->local_normal = poly.normals[rf.i1];
    m->local_point = poly.vertices[rf.i1];
  }
  int pc = 0;
  for (int i = 0; i < kMaxManifoldPoints; ++i) {
    real separation = dot(rf.normal, c2[i].v - rf.v1);
    if (separation <= radius) {
      ManifoldPoint* cp = &m->points[pc];
      if (primary.type == kEdgeA) {
        cp->local_point = mulT(xf, c2[i].v);
        cp->id = c2[i].id;
      } else {
        cp->local_point = c2[i].v;
        cp->id.type_a = c2[i].id.type_b;
        cp->id.type_b = c2[i].id.type_a;
        cp->id.index_a = c2[i].id.index_b;
        cp->id.index_b = c2[i].id.index_a;
      }
      ++pc;
    }
  }
  m->point_count = pc;
}

// Contact type registry: src/private/dynamics/b2_contact_registers.rs:67-103.
// Returns true when (typeA, typeB) is a registered primary pair; *swap is set when the
// reference creates the contact with A and B exchanged. Returns false when no contact
// class exists for the pair (edge-edge and chain pairs other than chain-circle/polygon).
inline bool registry_lookup(int typeA, int typeB, bool* swap) {
  auto primary = [](int a, int b) {
    return (a == kCircle && b == kCircle) || (a == kPolygon && b == kCircle) || (a == kPolygon && b == kPolygon) ||
           (a == kEdge && b == kCircle) || (a == kEdge && b == kPolygon) || (a == kChain && b == kCircle) ||
           (a == kChain && b == kPolygon);
  };
  if (primary(typeA, typeB)) { *swap = false; return true; }
  if (primary(typeB, typeA)) { *swap = true; return true; }
  return false;
}

// B2contactDynTrait::evaluate dispatch (src/private/dynamics/contacts/*.rs:45-51): the shapes
// must already be in registry order (A primary).
inline void evaluate(Manifold* m, const Shape& a, const Transform& xfA, const Shape& b, const Transform& xfB) {
  if (a.type == kCircle && b.type == kCircle) collide_circles(m, a, xfA, b, xfB);
  else if (a.type == kPolygon && b.type == kCircle) collide_polygon_circle(m, a, xfA, b, xfB);
  else if (a.type == kPolygon && b.type == kPolygon) collide_polygons(m, a, xfA, b, xfB);
  else if (a.type == kEdge && b.type == kCircle) collide_edge_circle(m, a, xfA, b, xfB);
  else if (a.type == kEdge && b.type == kPolygon) collide_edge_polygon(m, a, xfA, b, xfB);
  else m->point_count = 0;
}

}  // namespace b2o
