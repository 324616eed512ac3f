// b2g: narrowphase collision for one candidate pair, run by one GPU thread.
//
// The five manifold generators of the reference (box2d-rs 0.0.4) rewritten for registers:
//   circles            src/private/collision/b2_collide_circle.rs:7-34
//   polygon vs circle  src/private/collision/b2_collide_circle.rs:36-133
//   polygons           src/private/collision/b2_collide_polygon.rs:8-225
//   edge vs circle     src/private/collision/b2_collide_edge.rs:11-123
//   edge vs polygon    src/private/collision/b2_collide_edge.rs:125-475
// plus clipping (src/private/collision/b2_collision.rs:171-214) and the world manifold
// (src/private/collision/b2_collision.rs:7-68). Contact feature ids are packed into one
// uint32 (index_a | index_b << 8 | type_a << 16 | type_b << 24) so that warm-start matching
// is a single integer compare; the packing is the byte order of B2contactFeature.
//
// The generators are templates over the scalar T of the transforms and of the manifold
// geometry (float in the kernels; double and Dual for the adjoint), while shapes stay float.
// Branches compare primal values, so a Dual evaluation follows the float path's features.
#pragma once
#include "b2g_math.h"

namespace b2g {

enum ShapeType : int { kCircle = 0, kEdge = 1, kPolygon = 2 };
enum ManifoldType : int { kCircles = 0, kFaceA = 1, kFaceB = 2 };
enum FeatureType : uint32_t { kVertex = 0, kFace = 1 };

B2G_HD uint32_t feature(uint32_t ia, uint32_t ib, uint32_t ta, uint32_t tb) {
  return (ia & 0xffu) | (ib & 0xffu) << 8 | ta << 16 | tb << 24;
}
B2G_HD uint32_t feature_swap(uint32_t id) {
  return ((id & 0xffu) << 8) | ((id >> 8) & 0xffu) | (((id >> 16) & 0xffu) << 24) | (((id >> 24) & 0xffu) << 16);
}

// A shape in registers. Polygon: v/n hold vertices and normals, p the centroid. Circle: p is
// the centre. Edge: v[0]=vertex1, v[1]=vertex2, v[2]=vertex0, v[3]=vertex3.
static constexpr int kMaxPolygonVertices = 8;  // B2_MAX_POLYGON_VERTICES, src/b2_settings.rs

struct ShapeV {
  int type;
  int count;
  int one_sided;
  float radius;
  V2 p;
  V2 v[kMaxPolygonVertices];
  V2 n[kMaxPolygonVertices];
};

template <class T>
struct TManifold {
  int type;
  int count;
  TV2<T> local_normal;
  TV2<T> local_point;
  TV2<T> lp[2];
  uint32_t id[2];
};
using Manifold = TManifold<float>;

template <class T>
B2G_HD void manifold_clear(TManifold<T>* m) {
  m->type = kCircles;
  m->count = 0;
  m->local_normal = v2(T(0.0), T(0.0));
  m->local_point = v2(T(0.0), T(0.0));
  m->lp[0] = v2(T(0.0), T(0.0));
  m->lp[1] = v2(T(0.0), T(0.0));
  m->id[0] = 0;
  m->id[1] = 0;
}

template <class T>
struct TClipV {
  TV2<T> v;
  uint32_t id;
};

// b2_clip_segment_to_line (src/private/collision/b2_collision.rs): the points behind the plane
// in input order, then the crossing point when the two lie on opposite sides. Written with
// selects only (no run-time index into a local array, which would live in scratch memory); when
// both points are kept the product d0 * d1 is not negative, so at most two points come out.
template <class T>
B2G_HD int clip_segment(TClipV<T> out[2], const TClipV<T> in[2], TV2<T> normal, T offset, int vertexIndexA) {
  const T d0 = dot(normal, in[0].v) - offset;
  const T d1 = dot(normal, in[1].v) - offset;
  const bool keep0 = d0 <= T(0.0), keep1 = d1 <= T(0.0), crossing = d0 * d1 < T(0.0);
  TClipV<T> mid = in[0];
  if (crossing) {
    T interp = d0 / (d0 - d1);
    mid.v = in[0].v + interp * (in[1].v - in[0].v);
    mid.id = feature(uint32_t(vertexIndexA), (in[0].id >> 8) & 0xffu, kVertex, kFace);
  }
  const int kept = int(keep0) + int(keep1);
  const TClipV<T> first = keep0 ? in[0] : in[1];
  out[0] = kept > 0 ? first : mid;
  out[1] = kept > 1 ? in[1] : mid;
  return kept + int(crossing);
}

// manifold point pc (0 or 1) without a run-time index into the manifold: both slots are stored
// with selected values (a branch per slot would be merged into a store to a selected address)
template <class T>
B2G_HD void set_point(TManifold<T>* m, int pc, TV2<T> p, uint32_t id) {
  const bool first = pc == 0;
  m->lp[0] = first ? p : m->lp[0];
  m->id[0] = first ? id : m->id[0];
  m->lp[1] = first ? m->lp[1] : p;
  m->id[1] = first ? m->id[1] : id;
}

template <class T>
B2G_HD void collide_circles(TManifold<T>* m, const ShapeV& a, TXf<T> xfA, const ShapeV& b, TXf<T> xfB) {
  TV2<T> pA = mul(xfA, tv<T>(a.p)), pB = mul(xfB, tv<T>(b.p));
  TV2<T> d = pB - pA;
  T distSqr = dot(d, d);
  T radius = T(a.radius) + T(b.radius);
  if (distSqr > radius * radius) return;
  m->type = kCircles;
  m->local_point = tv<T>(a.p);
  m->local_normal = v2(T(0.0), T(0.0));
  m->count = 1;
  m->lp[0] = tv<T>(b.p);
  m->id[0] = 0;
}

template <class T>
B2G_HD void collide_polygon_circle(TManifold<T>* m, const ShapeV& poly, TXf<T> xfA, const ShapeV& circ, TXf<T> xfB) {
  TV2<T> c = mul(xfB, tv<T>(circ.p));
  TV2<T> cl = mulT(xfA, c);
  int ni = 0;
  T separation = T(-kMaxFloat);
  T radius = T(poly.radius) + T(circ.radius);
  const int vc = poly.count;
  TV2<T> pn[kMaxPolygonVertices], pv[kMaxPolygonVertices];  // read once, fixed indices after unrolling
#pragma unroll
  for (int i = 0; i < kMaxPolygonVertices; ++i) {
    pn[i] = i < vc ? tv<T>(poly.n[i]) : v2(T(0.0), T(0.0));
    pv[i] = i < vc ? tv<T>(poly.v[i]) : v2(T(0.0), T(0.0));
  }
#pragma unroll
  for (int i = 0; i < kMaxPolygonVertices; ++i) {
    if (i >= vc) break;
    T s = dot(pn[i], cl - pv[i]);
    if (s > radius) return;
    if (s > separation) { separation = s; ni = i; }
  }
  int i1 = ni, i2 = i1 + 1 < vc ? i1 + 1 : 0;
  TV2<T> v1 = tv<T>(poly.v[i1]), v2_ = tv<T>(poly.v[i2]);
  if (separation < T(kEpsilon)) {
    m->count = 1; m->type = kFaceA;
    m->local_normal = tv<T>(poly.n[ni]);
    m->local_point = T(0.5) * (v1 + v2_);
    m->lp[0] = tv<T>(circ.p); m->id[0] = 0;
    return;
  }
  T u1 = dot(cl - v1, v2_ - v1);
  T u2 = dot(cl - v2_, v1 - v2_);
  if (u1 <= T(0.0)) {
    if (dist_sq(cl, v1) > radius * radius) return;
    m->count = 1; m->type = kFaceA;
    TV2<T> n = cl - v1; normalize(&n);
    m->local_normal = n; m->local_point = v1;
    m->lp[0] = tv<T>(circ.p); m->id[0] = 0;
  } else if (u2 <= T(0.0)) {
    if (dist_sq(cl, v2_) > radius * radius) return;
    m->count = 1; m->type = kFaceA;
    TV2<T> n = cl - v2_; normalize(&n);
    m->local_normal = n; m->local_point = v2_;
    m->lp[0] = tv<T>(circ.p); m->id[0] = 0;
  } else {
    TV2<T> fc = T(0.5) * (v1 + v2_);
    T s = dot(cl - fc, tv<T>(poly.n[i1]));
    if (s > radius) return;
    m->count = 1; m->type = kFaceA;
    m->local_normal = tv<T>(poly.n[i1]); m->local_point = fc;
    m->lp[0] = tv<T>(circ.p); m->id[0] = 0;
  }
}

template <class T>
B2G_HD T find_max_separation(int* edgeIndex, const ShapeV& p1, TXf<T> xf1, const ShapeV& p2, TXf<T> xf2) {
  TXf<T> xf = mulT(xf2, xf1);
  int best = 0;
  T maxSep = T(-kMaxFloat);
  // p2's vertices are read once into registers (fixed indices after unrolling)
  const int n2 = p2.count;
  TV2<T> w2[kMaxPolygonVertices];
#pragma unroll
  for (int j = 0; j < kMaxPolygonVertices; ++j) w2[j] = j < n2 ? tv<T>(p2.v[j]) : v2(T(0.0), T(0.0));
  const int n1 = p1.count;
  TV2<T> pn[kMaxPolygonVertices], pv[kMaxPolygonVertices];
#pragma unroll
  for (int i = 0; i < kMaxPolygonVertices; ++i) {
    pn[i] = i < n1 ? tv<T>(p1.n[i]) : v2(T(0.0), T(0.0));
    pv[i] = i < n1 ? tv<T>(p1.v[i]) : v2(T(0.0), T(0.0));
  }
#pragma unroll
  for (int i = 0; i < kMaxPolygonVertices; ++i) {
    if (i >= n1) break;
    TV2<T> n = mul(xf.q, pn[i]);
    TV2<T> v1 = mul(xf, pv[i]);
    T si = T(kMaxFloat);
#pragma unroll
    for (int j = 0; j < kMaxPolygonVertices; ++j) {
      if (j >= n2) break;
      T sij = dot(n, w2[j] - v1);
      if (sij < si) si = sij;
    }
    if (si > maxSep) { maxSep = si; best = i; }
  }
  *edgeIndex = best;
  return maxSep;
}

template <class T>
B2G_HD void collide_polygons(TManifold<T>* m, const ShapeV& pA, TXf<T> xfA, const ShapeV& pB, TXf<T> xfB) {
  T totalRadius = T(pA.radius) + T(pB.radius);
  int edgeA = 0;
  T sepA = find_max_separation(&edgeA, pA, xfA, pB, xfB);
  if (sepA > totalRadius) return;
  int edgeB = 0;
  T sepB = find_max_separation(&edgeB, pB, xfB, pA, xfA);
  if (sepB > totalRadius) return;
  const T kTol = T(0.1) * T(kLinearSlopD);
  const bool flip = sepB > sepA + kTol;
  const ShapeV& poly1 = flip ? pB : pA;
  const ShapeV& poly2 = flip ? pA : pB;
  TXf<T> xf1 = flip ? xfB : xfA;
  TXf<T> xf2 = flip ? xfA : xfB;
  int edge1 = flip ? edgeB : edgeA;
  int mtype = flip ? kFaceB : kFaceA;
  // incident edge (src/private/collision/b2_collide_polygon.rs:48-93)
  TV2<T> normal1 = mulT(xf2.q, mul(xf1.q, tv<T>(poly1.n[edge1])));
  int index = 0;
  T minDot = T(kMaxFloat);
  for (int i = 0; i < poly2.count; ++i) {
    T d = dot(normal1, tv<T>(poly2.n[i]));
    if (d < minDot) { minDot = d; index = i; }
  }
  int j1 = index, j2 = j1 + 1 < poly2.count ? j1 + 1 : 0;
  TClipV<T> incident[2];
  incident[0].v = mul(xf2, tv<T>(poly2.v[j1]));
  incident[0].id = feature(uint32_t(edge1), uint32_t(j1), kFace, kVertex);
  incident[1].v = mul(xf2, tv<T>(poly2.v[j2]));
  incident[1].id = feature(uint32_t(edge1), uint32_t(j2), kFace, kVertex);
  int count1 = poly1.count;
  int iv1 = edge1, iv2 = edge1 + 1 < count1 ? edge1 + 1 : 0;
  TV2<T> v11 = tv<T>(poly1.v[iv1]), v12 = tv<T>(poly1.v[iv2]);
  TV2<T> localTangent = v12 - v11;
  normalize(&localTangent);
  TV2<T> localNormal = cross(localTangent, T(1.0));
  TV2<T> planePoint = T(0.5) * (v11 + v12);
  TV2<T> tangent = mul(xf1.q, localTangent);
  TV2<T> normal = cross(tangent, T(1.0));
  v11 = mul(xf1, v11);
  v12 = mul(xf1, v12);
  T frontOffset = dot(normal, v11);
  T sideOffset1 = -dot(tangent, v11) + totalRadius;
  T sideOffset2 = dot(tangent, v12) + totalRadius;
  TClipV<T> cp1[2], cp2[2];
  int np = clip_segment(cp1, incident, -tangent, sideOffset1, iv1);
  if (np < 2) { m->type = mtype; return; }
  np = clip_segment(cp2, cp1, tangent, sideOffset2, iv2);
  if (np < 2) { m->type = mtype; return; }
  m->type = mtype;
  m->local_normal = localNormal;
  m->local_point = planePoint;
  int pc = 0;
  for (int i = 0; i < 2; ++i) {
    T separation = dot(normal, cp2[i].v) - frontOffset;
    if (separation <= totalRadius) {
      set_point(m, pc, mulT(xf2, cp2[i].v), flip ? feature_swap(cp2[i].id) : cp2[i].id);
      ++pc;
    }
  }
  m->count = pc;
}

template <class T>
B2G_HD void collide_edge_circle(TManifold<T>* m, const ShapeV& edge, TXf<T> xfA, const ShapeV& circ, TXf<T> xfB) {
  TV2<T> Q = mulT(xfA, mul(xfB, tv<T>(circ.p)));
  TV2<T> A = tv<T>(edge.v[0]), B = tv<T>(edge.v[1]);
  TV2<T> e = B - A;
  TV2<T> n = v2(e.y, -e.x);
  T offset = dot(n, Q - A);
  bool oneSided = edge.one_sided != 0;
  if (oneSided && offset < T(0.0)) return;
  T u = dot(e, B - Q);
  T v = dot(e, Q - A);
  T radius = T(edge.radius) + T(circ.radius);
  if (v <= T(0.0)) {
    TV2<T> d = Q - A;
    if (dot(d, d) > radius * radius) return;
    if (oneSided) {
      TV2<T> e1 = A - tv<T>(edge.v[2]);
      if (dot(e1, A - Q) > T(0.0)) return;
    }
    m->count = 1; m->type = kCircles; m->local_normal = v2(T(0.0), T(0.0)); m->local_point = A;
    m->id[0] = feature(0, 0, kVertex, kVertex); m->lp[0] = tv<T>(circ.p);
    return;
  }
  if (u <= T(0.0)) {
    TV2<T> d = Q - B;
    if (dot(d, d) > radius * radius) return;
    if (oneSided) {
      TV2<T> e2 = tv<T>(edge.v[3]) - B;
      if (dot(e2, Q - B) > T(0.0)) return;
    }
    m->count = 1; m->type = kCircles; m->local_normal = v2(T(0.0), T(0.0)); m->local_point = B;
    m->id[0] = feature(1, 0, kVertex, kVertex); m->lp[0] = tv<T>(circ.p);
    return;
  }
  T den = dot(e, e);
  TV2<T> P = (T(1.0) / den) * (u * A + v * B);
  TV2<T> d = Q - P;
  if (dot(d, d) > radius * radius) return;
  if (offset < T(0.0)) n = v2(-n.x, -n.y);
  normalize(&n);
  m->count = 1; m->type = kFaceA; m->local_normal = n; m->local_point = A;
  m->id[0] = feature(0, 0, kFace, kVertex); m->lp[0] = tv<T>(circ.p);
}

template <class T>
B2G_HD void collide_edge_polygon(TManifold<T>* m, const ShapeV& edge, TXf<T> xfA, const ShapeV& poly, TXf<T> xfB) {
  TXf<T> xf = mulT(xfA, xfB);
  TV2<T> centroidB = mul(xf, tv<T>(poly.p));
  TV2<T> v1 = tv<T>(edge.v[0]), v2_ = tv<T>(edge.v[1]);
  TV2<T> edge1 = v2_ - v1;
  normalize(&edge1);
  TV2<T> normal1 = v2(edge1.y, -edge1.x);
  T offset1 = dot(normal1, centroidB - v1);
  bool oneSided = edge.one_sided != 0;
  if (oneSided && offset1 < T(0.0)) return;
  // the polygon in the edge's frame, recomputed where used (a local array indexed at run time
  // would live in scratch memory)
  const int tc = poly.count;
  auto tvx = [&](int i) { return mul(xf, tv<T>(poly.v[i])); };
  auto tn = [&](int i) { return mul(xf.q, tv<T>(poly.n[i])); };
  T radius = T(poly.radius) + T(edge.radius);
  // edge separation (src/private/collision/b2_collide_edge.rs:172-201)
  int eIndex = -1;
  T eSep = T(-kMaxFloat);
  TV2<T> eNormal = v2(T(0.0), T(0.0));
  {
    TV2<T> axes[2] = {normal1, -normal1};
    for (int j = 0; j < 2; ++j) {
      T sj = T(kMaxFloat);
      for (int i = 0; i < tc; ++i) {
        T si = dot(axes[j], tvx(i) - v1);
        if (si < sj) sj = si;
      }
      if (sj > eSep) { eIndex = j; eSep = sj; eNormal = axes[j]; }
    }
  }
  if (eSep > radius) return;
  // polygon separation (src/private/collision/b2_collide_edge.rs:203-228)
  int pType = 0, pIndex = -1;
  T pSep = T(-kMaxFloat);
  TV2<T> pNormal = v2(T(0.0), T(0.0));
  for (int i = 0; i < tc; ++i) {
    TV2<T> n = -tn(i);
    T s1 = dot(n, tvx(i) - v1);
    T s2 = dot(n, tvx(i) - v2_);
    T s = bmin(s1, s2);
    if (s > pSep) { pType = 2; pIndex = i; pSep = s; pNormal = n; }
  }
  if (pSep > radius) return;
  const T kRelativeTol = T(0.98), kAbsoluteTol = T(0.001);
  // primary axis: type 1 = edge A, 2 = edge B, 0 = unknown
  int prType, prIndex;
  TV2<T> prNormal;
  if (pSep - radius > kRelativeTol * (eSep - radius) + kAbsoluteTol) {
    prType = pType; prIndex = pIndex; prNormal = pNormal;
  } else {
    prType = 1; prIndex = eIndex; prNormal = eNormal;
  }
  if (oneSided) {
    TV2<T> edge0 = v1 - tv<T>(edge.v[2]);
    normalize(&edge0);
    TV2<T> normal0 = v2(edge0.y, -edge0.x);
    bool convex1 = cross(edge0, edge1) >= T(0.0);
    TV2<T> edge2 = tv<T>(edge.v[3]) - v2_;
    normalize(&edge2);
    TV2<T> normal2 = v2(edge2.y, -edge2.x);
    bool convex2 = cross(edge1, edge2) >= T(0.0);
    const T sinTol = T(0.1);
    bool side1 = dot(prNormal, edge1) <= T(0.0);
    if (side1) {
      if (convex1) {
        if (cross(prNormal, normal0) > sinTol) return;
      } else {
        prType = 1; prIndex = eIndex; prNormal = eNormal;
      }
    } else {
      if (convex2) {
        if (cross(normal2, prNormal) > sinTol) return;
      } else {
        prType = 1; prIndex = eIndex; prNormal = eNormal;
      }
    }
  }
  TClipV<T> clip[2];
  int ri1, ri2;
  TV2<T> rv1, rv2, rnormal, sn1, sn2;
  if (prType == 1) {
    m->type = kFaceA;
    int bestIndex = 0;
    T bestValue = dot(prNormal, tn(0));
    for (int i = 1; i < tc; ++i) {
      T value = dot(prNormal, tn(i));
      if (value < bestValue) { bestValue = value; bestIndex = i; }
    }
    int i1 = bestIndex, i2 = i1 + 1 < tc ? i1 + 1 : 0;
    clip[0].v = tvx(i1);
    clip[0].id = feature(0, uint32_t(i1), kFace, kVertex);
    clip[1].v = tvx(i2);
    clip[1].id = feature(0, uint32_t(i2), kFace, kVertex);
    ri1 = 0; ri2 = 1; rv1 = v1; rv2 = v2_; rnormal = prNormal;
    sn1 = -edge1; sn2 = edge1;
  } else {
    m->type = kFaceB;
    clip[0].v = v2_;
    clip[0].id = feature(1, uint32_t(prIndex), kVertex, kFace);
    clip[1].v = v1;
    clip[1].id = feature(0, uint32_t(prIndex), kVertex, kFace);
    ri1 = prIndex;
    ri2 = ri1 + 1 < tc ? ri1 + 1 : 0;
    rv1 = tvx(ri1); rv2 = tvx(ri2);
    rnormal = tn(ri1);
    sn1 = v2(rnormal.y, -rnormal.x);
    sn2 = -sn1;
  }
  T so1 = dot(sn1, rv1);
  T so2 = dot(sn2, rv2);
  TClipV<T> c1[2], c2[2];
  int np = clip_segment(c1, clip, sn1, so1, ri1);
  if (np < 2) return;
  np = clip_segment(c2, c1, sn2, so2, ri2);
  if (np < 2) return;
  if (prType == 1) {
    m->local_normal = rnormal;
    m->local_point = rv1;
  } else {
    m->local_normal = tv<T>(poly.n[ri1]);
    m->local_point = tv<T>(poly.v[ri1]);
  }
  int pc = 0;
  for (int i = 0; i < 2; ++i) {
    T separation = dot(rnormal, c2[i].v - rv1);
    if (separation <= radius) {
      if (prType == 1) set_point(m, pc, mulT(xf, c2[i].v), c2[i].id);
      else set_point(m, pc, c2[i].v, feature_swap(c2[i].id));
      ++pc;
    }
  }
  m->count = pc;
}

// evaluate() for a pair already in registry order (A is the primary shape type).
template <class T>
B2G_HD void evaluate(TManifold<T>* m, const ShapeV& a, TXf<T> xfA, const ShapeV& b, TXf<T> xfB) {
  manifold_clear(m);
  if (a.type == kCircle) {
    collide_circles(m, a, xfA, b, xfB);
  } else if (a.type == kPolygon) {
    if (b.type == kCircle) collide_polygon_circle(m, a, xfA, b, xfB);
    else collide_polygons(m, a, xfA, b, xfB);
  } else {
    if (b.type == kCircle) collide_edge_circle(m, a, xfA, b, xfB);
    else collide_edge_polygon(m, a, xfA, b, xfB);
  }
  if (m->count == 0) manifold_clear(m);
}

// Registry order (src/private/dynamics/b2_contact_registers.rs:67-103): returns 1 when the
// reference creates the contact with A/B exchanged, 0 when as given, -1 when no contact class
// exists for the pair of shape types.
B2G_HD int registry_swap(int ta, int tb) {
  // primary pairs: (c,c) (p,c) (p,p) (e,c) (e,p)
  auto prim = [](int a, int b) {
    return (a == kCircle && b == kCircle) || (a == kPolygon && b == kCircle) || (a == kPolygon && b == kPolygon) ||
           (a == kEdge && b == kCircle) || (a == kEdge && b == kPolygon);
  };
  if (prim(ta, tb)) return 0;
  if (prim(tb, ta)) return 1;
  return -1;
}

// World manifold: normal, points, separations (src/private/collision/b2_collision.rs:7-68).
template <class T>
struct TWorldManifold {
  TV2<T> normal;
  TV2<T> points[2];
  T sep[2];
};
using WorldManifold = TWorldManifold<float>;

template <class T>
B2G_HD void world_manifold(TWorldManifold<T>* wm, int type, int count, TV2<T> ln, TV2<T> lpt, const TV2<T>* lp,
                           TXf<T> xfA, T rA, TXf<T> xfB, T rB) {
  if (count == 0) return;
  if (type == kCircles) {
    wm->normal = v2(T(1.0), T(0.0));
    TV2<T> pA = mul(xfA, lpt);
    TV2<T> pB = mul(xfB, lp[0]);
    if (dist_sq(pA, pB) > T(kEpsilon) * T(kEpsilon)) {
      wm->normal = pB - pA;
      normalize(&wm->normal);
    }
    TV2<T> cA = pA + rA * wm->normal;
    TV2<T> cB = pB - rB * wm->normal;
    wm->points[0] = T(0.5) * (cA + cB);
    wm->sep[0] = dot(cB - cA, wm->normal);
  } else if (type == kFaceA) {
    wm->normal = mul(xfA.q, ln);
    TV2<T> plane = mul(xfA, lpt);
    for (int i = 0; i < count; ++i) {
      TV2<T> clip = mul(xfB, lp[i]);
      TV2<T> cA = clip + (rA - dot(clip - plane, wm->normal)) * wm->normal;
      TV2<T> cB = clip - rB * wm->normal;
      wm->points[i] = T(0.5) * (cA + cB);
      wm->sep[i] = dot(cB - cA, wm->normal);
    }
  } else {
    wm->normal = mul(xfB.q, ln);
    TV2<T> plane = mul(xfB, lpt);
    for (int i = 0; i < count; ++i) {
      TV2<T> clip = mul(xfA, lp[i]);
      TV2<T> cB = clip + (rB - dot(clip - plane, wm->normal)) * wm->normal;
      TV2<T> cA = clip - rA * wm->normal;
      wm->points[i] = T(0.5) * (cA + cB);
      wm->sep[i] = dot(cA - cB, wm->normal);
    }
    wm->normal = -wm->normal;
  }
}

// Shape AABB under a transform (compute_aabb of the three shape types).
B2G_HD Aabb shape_aabb(const ShapeV& s, Xf xf) {
  Aabb a;
  if (s.type == kCircle) {
    V2 p = xf.p + mul(xf.q, s.p);
    a.lo = v2(p.x - s.radius, p.y - s.radius);
    a.hi = v2(p.x + s.radius, p.y + s.radius);
  } else if (s.type == kPolygon) {
    V2 lo = mul(xf, s.v[0]);
    V2 hi = lo;
    for (int i = 1; i < s.count; ++i) {
      V2 v = mul(xf, s.v[i]);
      lo = vmin(lo, v);
      hi = vmax(hi, v);
    }
    V2 r = v2(s.radius, s.radius);
    a.lo = lo - r;
    a.hi = hi + r;
  } else {
    V2 p1 = mul(xf, s.v[0]), p2 = mul(xf, s.v[1]);
    V2 r = v2(s.radius, s.radius);
    a.lo = vmin(p1, p2) - r;
    a.hi = vmax(p1, p2) + r;
  }
  return a;
}

}  // namespace b2g
