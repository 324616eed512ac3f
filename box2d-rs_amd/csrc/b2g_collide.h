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
struct ShapeV {
  int type;
  int count;
  int one_sided;
  float radius;
  V2 p;
  V2 v[8];
  V2 n[8];
};

struct Manifold {
  int type;
  int count;
  V2 local_normal;
  V2 local_point;
  V2 lp[2];
  uint32_t id[2];
};

B2G_HD void manifold_clear(Manifold* m) {
  m->type = kCircles;
  m->count = 0;
  m->local_normal = v2(0.0f, 0.0f);
  m->local_point = v2(0.0f, 0.0f);
  m->lp[0] = v2(0.0f, 0.0f);
  m->lp[1] = v2(0.0f, 0.0f);
  m->id[0] = 0;
  m->id[1] = 0;
}

struct ClipV {
  V2 v;
  uint32_t id;
};

B2G_HD int clip_segment(ClipV out[2], const ClipV in[2], V2 normal, float offset, int vertexIndexA) {
  int count = 0;
  float d0 = dot(normal, in[0].v) - offset;
  float d1 = dot(normal, in[1].v) - offset;
  if (d0 <= 0.0f) out[count++] = in[0];
  if (d1 <= 0.0f) out[count++] = in[1];
  if (d0 * d1 < 0.0f) {
    float interp = d0 / (d0 - d1);
    out[count].v = in[0].v + interp * (in[1].v - in[0].v);
    out[count].id = feature(uint32_t(vertexIndexA), (in[0].id >> 8) & 0xffu, kVertex, kFace);
    ++count;
  }
  return count;
}

B2G_HD void collide_circles(Manifold* m, const ShapeV& a, Xf xfA, const ShapeV& b, Xf xfB) {
  V2 pA = mul(xfA, a.p), pB = mul(xfB, b.p);
  V2 d = pB - pA;
  float distSqr = dot(d, d);
  float radius = a.radius + b.radius;
  if (distSqr > radius * radius) return;
  m->type = kCircles;
  m->local_point = a.p;
  m->local_normal = v2(0.0f, 0.0f);
  m->count = 1;
  m->lp[0] = b.p;
  m->id[0] = 0;
}

B2G_HD void collide_polygon_circle(Manifold* m, const ShapeV& poly, Xf xfA, const ShapeV& circ, Xf xfB) {
  V2 c = mul(xfB, circ.p);
  V2 cl = mulT(xfA, c);
  int ni = 0;
  float separation = -kMaxFloat;
  float radius = poly.radius + circ.radius;
  int vc = poly.count;
  for (int i = 0; i < vc; ++i) {
    float s = dot(poly.n[i], cl - poly.v[i]);
    if (s > radius) return;
    if (s > separation) { separation = s; ni = i; }
  }
  int i1 = ni, i2 = i1 + 1 < vc ? i1 + 1 : 0;
  V2 v1 = poly.v[i1], v2_ = poly.v[i2];
  if (separation < kEpsilon) {
    m->count = 1; m->type = kFaceA;
    m->local_normal = poly.n[ni];
    m->local_point = 0.5f * (v1 + v2_);
    m->lp[0] = circ.p; m->id[0] = 0;
    return;
  }
  float u1 = dot(cl - v1, v2_ - v1);
  float u2 = dot(cl - v2_, v1 - v2_);
  if (u1 <= 0.0f) {
    if (dist_sq(cl, v1) > radius * radius) return;
    m->count = 1; m->type = kFaceA;
    V2 n = cl - v1; normalize(&n);
    m->local_normal = n; m->local_point = v1;
    m->lp[0] = circ.p; m->id[0] = 0;
  } else if (u2 <= 0.0f) {
    if (dist_sq(cl, v2_) > radius * radius) return;
    m->count = 1; m->type = kFaceA;
    V2 n = cl - v2_; normalize(&n);
    m->local_normal = n; m->local_point = v2_;
    m->lp[0] = circ.p; m->id[0] = 0;
  } else {
    V2 fc = 0.5f * (v1 + v2_);
    float s = dot(cl - fc, poly.n[i1]);
    if (s > radius) return;
    m->count = 1; m->type = kFaceA;
    m->local_normal = poly.n[i1]; m->local_point = fc;
    m->lp[0] = circ.p; m->id[0] = 0;
  }
}

B2G_HD float find_max_separation(int* edgeIndex, const ShapeV& p1, Xf xf1, const ShapeV& p2, Xf xf2) {
  Xf xf = mulT(xf2, xf1);
  int best = 0;
  float maxSep = -kMaxFloat;
  for (int i = 0; i < p1.count; ++i) {
    V2 n = mul(xf.q, p1.n[i]);
    V2 v1 = mul(xf, p1.v[i]);
    float si = kMaxFloat;
    for (int j = 0; j < p2.count; ++j) {
      float sij = dot(n, p2.v[j] - v1);
      if (sij < si) si = sij;
    }
    if (si > maxSep) { maxSep = si; best = i; }
  }
  *edgeIndex = best;
  return maxSep;
}

B2G_HD void collide_polygons(Manifold* m, const ShapeV& pA, Xf xfA, const ShapeV& pB, Xf xfB) {
  float totalRadius = pA.radius + pB.radius;
  int edgeA = 0;
  float sepA = find_max_separation(&edgeA, pA, xfA, pB, xfB);
  if (sepA > totalRadius) return;
  int edgeB = 0;
  float sepB = find_max_separation(&edgeB, pB, xfB, pA, xfA);
  if (sepB > totalRadius) return;
  const float kTol = 0.1f * kLinearSlop;
  const bool flip = sepB > sepA + kTol;
  const ShapeV& poly1 = flip ? pB : pA;
  const ShapeV& poly2 = flip ? pA : pB;
  Xf xf1 = flip ? xfB : xfA;
  Xf xf2 = flip ? xfA : xfB;
  int edge1 = flip ? edgeB : edgeA;
  int mtype = flip ? kFaceB : kFaceA;
  // incident edge (src/private/collision/b2_collide_polygon.rs:48-93)
  V2 normal1 = mulT(xf2.q, mul(xf1.q, poly1.n[edge1]));
  int index = 0;
  float minDot = kMaxFloat;
  for (int i = 0; i < poly2.count; ++i) {
    float d = dot(normal1, poly2.n[i]);
    if (d < minDot) { minDot = d; index = i; }
  }
  int j1 = index, j2 = j1 + 1 < poly2.count ? j1 + 1 : 0;
  ClipV incident[2];
  incident[0].v = mul(xf2, poly2.v[j1]);
  incident[0].id = feature(uint32_t(edge1), uint32_t(j1), kFace, kVertex);
  incident[1].v = mul(xf2, poly2.v[j2]);
  incident[1].id = feature(uint32_t(edge1), uint32_t(j2), kFace, kVertex);
  int count1 = poly1.count;
  int iv1 = edge1, iv2 = edge1 + 1 < count1 ? edge1 + 1 : 0;
  V2 v11 = poly1.v[iv1], v12 = poly1.v[iv2];
  V2 localTangent = v12 - v11;
  normalize(&localTangent);
  V2 localNormal = cross(localTangent, 1.0f);
  V2 planePoint = 0.5f * (v11 + v12);
  V2 tangent = mul(xf1.q, localTangent);
  V2 normal = cross(tangent, 1.0f);
  v11 = mul(xf1, v11);
  v12 = mul(xf1, v12);
  float frontOffset = dot(normal, v11);
  float sideOffset1 = -dot(tangent, v11) + totalRadius;
  float sideOffset2 = dot(tangent, v12) + totalRadius;
  ClipV cp1[2], cp2[2];
  int np = clip_segment(cp1, incident, -tangent, sideOffset1, iv1);
  if (np < 2) { m->type = mtype; return; }
  np = clip_segment(cp2, cp1, tangent, sideOffset2, iv2);
  if (np < 2) { m->type = mtype; return; }
  m->type = mtype;
  m->local_normal = localNormal;
  m->local_point = planePoint;
  int pc = 0;
  for (int i = 0; i < 2; ++i) {
    float separation = dot(normal, cp2[i].v) - frontOffset;
    if (separation <= totalRadius) {
      m->lp[pc] = mulT(xf2, cp2[i].v);
      m->id[pc] = flip ? feature_swap(cp2[i].id) : cp2[i].id;
      ++pc;
    }
  }
  m->count = pc;
}

B2G_HD void collide_edge_circle(Manifold* m, const ShapeV& edge, Xf xfA, const ShapeV& circ, Xf xfB) {
  V2 Q = mulT(xfA, mul(xfB, circ.p));
  V2 A = edge.v[0], B = edge.v[1];
  V2 e = B - A;
  V2 n = v2(e.y, -e.x);
  float offset = dot(n, Q - A);
  bool oneSided = edge.one_sided != 0;
  if (oneSided && offset < 0.0f) return;
  float u = dot(e, B - Q);
  float v = dot(e, Q - A);
  float radius = edge.radius + circ.radius;
  if (v <= 0.0f) {
    V2 d = Q - A;
    if (dot(d, d) > radius * radius) return;
    if (oneSided) {
      V2 e1 = A - edge.v[2];
      if (dot(e1, A - Q) > 0.0f) return;
    }
    m->count = 1; m->type = kCircles; m->local_normal = v2(0.0f, 0.0f); m->local_point = A;
    m->id[0] = feature(0, 0, kVertex, kVertex); m->lp[0] = circ.p;
    return;
  }
  if (u <= 0.0f) {
    V2 d = Q - B;
    if (dot(d, d) > radius * radius) return;
    if (oneSided) {
      V2 e2 = edge.v[3] - B;
      if (dot(e2, Q - B) > 0.0f) return;
    }
    m->count = 1; m->type = kCircles; m->local_normal = v2(0.0f, 0.0f); m->local_point = B;
    m->id[0] = feature(1, 0, kVertex, kVertex); m->lp[0] = circ.p;
    return;
  }
  float den = dot(e, e);
  V2 P = (1.0f / den) * (u * A + v * B);
  V2 d = Q - P;
  if (dot(d, d) > radius * radius) return;
  if (offset < 0.0f) n = v2(-n.x, -n.y);
  normalize(&n);
  m->count = 1; m->type = kFaceA; m->local_normal = n; m->local_point = A;
  m->id[0] = feature(0, 0, kFace, kVertex); m->lp[0] = circ.p;
}

B2G_HD void collide_edge_polygon(Manifold* m, const ShapeV& edge, Xf xfA, const ShapeV& poly, Xf xfB) {
  Xf xf = mulT(xfA, xfB);
  V2 centroidB = mul(xf, poly.p);
  V2 v1 = edge.v[0], v2_ = edge.v[1];
  V2 edge1 = v2_ - v1;
  normalize(&edge1);
  V2 normal1 = v2(edge1.y, -edge1.x);
  float offset1 = dot(normal1, centroidB - v1);
  bool oneSided = edge.one_sided != 0;
  if (oneSided && offset1 < 0.0f) return;
  V2 tv[8], tn[8];
  int tc = poly.count;
  for (int i = 0; i < tc; ++i) {
    tv[i] = mul(xf, poly.v[i]);
    tn[i] = mul(xf.q, poly.n[i]);
  }
  float radius = poly.radius + edge.radius;
  // edge separation (src/private/collision/b2_collide_edge.rs:172-201)
  int eIndex = -1;
  float eSep = -kMaxFloat;
  V2 eNormal = v2(0.0f, 0.0f);
  {
    V2 axes[2] = {normal1, -normal1};
    for (int j = 0; j < 2; ++j) {
      float sj = kMaxFloat;
      for (int i = 0; i < tc; ++i) {
        float si = dot(axes[j], tv[i] - v1);
        if (si < sj) sj = si;
      }
      if (sj > eSep) { eIndex = j; eSep = sj; eNormal = axes[j]; }
    }
  }
  if (eSep > radius) return;
  // polygon separation (src/private/collision/b2_collide_edge.rs:203-228)
  int pType = 0, pIndex = -1;
  float pSep = -kMaxFloat;
  V2 pNormal = v2(0.0f, 0.0f);
  for (int i = 0; i < tc; ++i) {
    V2 n = -tn[i];
    float s1 = dot(n, tv[i] - v1);
    float s2 = dot(n, tv[i] - v2_);
    float s = bmin(s1, s2);
    if (s > pSep) { pType = 2; pIndex = i; pSep = s; pNormal = n; }
  }
  if (pSep > radius) return;
  const float kRelativeTol = 0.98f, kAbsoluteTol = 0.001f;
  // primary axis: type 1 = edge A, 2 = edge B, 0 = unknown
  int prType, prIndex;
  V2 prNormal;
  if (pSep - radius > kRelativeTol * (eSep - radius) + kAbsoluteTol) {
    prType = pType; prIndex = pIndex; prNormal = pNormal;
  } else {
    prType = 1; prIndex = eIndex; prNormal = eNormal;
  }
  if (oneSided) {
    V2 edge0 = v1 - edge.v[2];
    normalize(&edge0);
    V2 normal0 = v2(edge0.y, -edge0.x);
    bool convex1 = cross(edge0, edge1) >= 0.0f;
    V2 edge2 = edge.v[3] - v2_;
    normalize(&edge2);
    V2 normal2 = v2(edge2.y, -edge2.x);
    bool convex2 = cross(edge1, edge2) >= 0.0f;
    const float sinTol = 0.1f;
    bool side1 = dot(prNormal, edge1) <= 0.0f;
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
  ClipV clip[2];
  int ri1, ri2;
  V2 rv1, rv2, rnormal, sn1, sn2;
  if (prType == 1) {
    m->type = kFaceA;
    int bestIndex = 0;
    float bestValue = dot(prNormal, tn[0]);
    for (int i = 1; i < tc; ++i) {
      float value = dot(prNormal, tn[i]);
      if (value < bestValue) { bestValue = value; bestIndex = i; }
    }
    int i1 = bestIndex, i2 = i1 + 1 < tc ? i1 + 1 : 0;
    clip[0].v = tv[i1];
    clip[0].id = feature(0, uint32_t(i1), kFace, kVertex);
    clip[1].v = tv[i2];
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
    rv1 = tv[ri1]; rv2 = tv[ri2];
    rnormal = tn[ri1];
    sn1 = v2(rnormal.y, -rnormal.x);
    sn2 = -sn1;
  }
  float so1 = dot(sn1, rv1);
  float so2 = dot(sn2, rv2);
  ClipV c1[2], c2[2];
  int np = clip_segment(c1, clip, sn1, so1, ri1);
  if (np < 2) return;
  np = clip_segment(c2, c1, sn2, so2, ri2);
  if (np < 2) return;
  if (prType == 1) {
    m->local_normal = rnormal;
    m->local_point = rv1;
  } else {
    m->local_normal = poly.n[ri1];
    m->local_point = poly.v[ri1];
  }
  int pc = 0;
  for (int i = 0; i < 2; ++i) {
    float separation = dot(rnormal, c2[i].v - rv1);
    if (separation <= radius) {
      if (prType == 1) {
        m->lp[pc] = mulT(xf, c2[i].v);
        m->id[pc] = c2[i].id;
      } else {
        m->lp[pc] = c2[i].v;
        m->id[pc] = feature_swap(c2[i].id);
      }
      ++pc;
    }
  }
  m->count = pc;
}

// evaluate() for a pair already in registry order (A is the primary shape type).
B2G_HD void evaluate(Manifold* m, const ShapeV& a, Xf xfA, const ShapeV& b, Xf xfB) {
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
struct WorldManifold {
  V2 normal;
  V2 points[2];
  float sep[2];
};

B2G_HD void world_manifold(WorldManifold* wm, int type, int count, V2 ln, V2 lpt, const V2* lp, Xf xfA, float rA,
                           Xf xfB, float rB) {
  if (count == 0) return;
  if (type == kCircles) {
    wm->normal = v2(1.0f, 0.0f);
    V2 pA = mul(xfA, lpt);
    V2 pB = mul(xfB, lp[0]);
    if (dist_sq(pA, pB) > kEpsilon * kEpsilon) {
      wm->normal = pB - pA;
      normalize(&wm->normal);
    }
    V2 cA = pA + rA * wm->normal;
    V2 cB = pB - rB * wm->normal;
    wm->points[0] = 0.5f * (cA + cB);
    wm->sep[0] = dot(cB - cA, wm->normal);
  } else if (type == kFaceA) {
    wm->normal = mul(xfA.q, ln);
    V2 plane = mul(xfA, lpt);
    for (int i = 0; i < count; ++i) {
      V2 clip = mul(xfB, lp[i]);
      V2 cA = clip + (rA - dot(clip - plane, wm->normal)) * wm->normal;
      V2 cB = clip - rB * wm->normal;
      wm->points[i] = 0.5f * (cA + cB);
      wm->sep[i] = dot(cB - cA, wm->normal);
    }
  } else {
    wm->normal = mul(xfB.q, ln);
    V2 plane = mul(xfB, lpt);
    for (int i = 0; i < count; ++i) {
      V2 clip = mul(xfA, lp[i]);
      V2 cB = clip + (rB - dot(clip - plane, wm->normal)) * wm->normal;
      V2 cA = clip - rA * wm->normal;
      wm->points[i] = 0.5f * (cA + cB);
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
