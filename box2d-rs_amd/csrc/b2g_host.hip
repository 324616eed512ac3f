// b2g host-side setup on the step path: shape construction and mass data.
//
// These run once per fixture on the host, in f32 with no FMA contraction, so that the shapes
// (vertex order, normals, centroid -- which fix the contact feature ids) and the mass data
// uploaded to the GPU are bit-identical to what the reference computes:
//   set_as_box / set_as_box_angle  src/private/collision/b2_polygon_shape.rs:12-52
//   set (gift-wrap hull)           src/private/collision/b2_polygon_shape.rs:96-211
//   centroid                       src/private/collision/b2_polygon_shape.rs:59-94
//   polygon mass                   src/private/collision/b2_polygon_shape.rs:314-390
//   circle mass                    src/private/collision/b2_circle_shape.rs:79-86
//   edge setters                   src/private/collision/b2_edge_shape.rs:7-19
#include <algorithm>

#include "b2g_internal.h"

namespace b2g {

static void host_box(HostShape* s, float hx, float hy) {
  s->type = kPolygon;
  s->radius = kPolygonRadius;
  s->count = 4;
  s->v[0] = v2(-hx, -hy);
  s->v[1] = v2(hx, -hy);
  s->v[2] = v2(hx, hy);
  s->v[3] = v2(-hx, hy);
  s->n[0] = v2(0.0f, -1.0f);
  s->n[1] = v2(1.0f, 0.0f);
  s->n[2] = v2(0.0f, 1.0f);
  s->n[3] = v2(-1.0f, 0.0f);
  s->p = v2(0.0f, 0.0f);
}

// Area-weighted centroid, triangles fanned from the first vertex.
static V2 host_centroid(const V2* vs, int count) {
  V2 c = v2(0.0f, 0.0f);
  float area = 0.0f;
  const V2 s = vs[0];
  const float inv3 = 1.0f / 3.0f;
  for (int i = 0; i < count; ++i) {
    V2 p1 = vs[0] - s;
    V2 p2 = vs[i] - s;
    V2 p3 = (i + 1 < count) ? vs[i + 1] - s : vs[0] - s;
    float d = cross(p2 - p1, p3 - p1);
    float tri = 0.5f * d;
    area += tri;
    c = c + (tri * inv3) * ((p1 + p2) + p3);
  }
  return (1.0f / area) * c + s;
}

int build_host_shape(const b2g_shape_def& d, HostShape* s) {
  *s = HostShape();
  switch (d.type) {
    case B2G_SHAPE_CIRCLE:
      s->type = kCircle;
      s->p = v2(d.params[0], d.params[1]);
      s->radius = d.radius;
      return B2G_OK;
    case B2G_SHAPE_EDGE:
      s->type = kEdge;
      s->radius = kPolygonRadius;
      s->v[0] = v2(d.params[0], d.params[1]);
      s->v[1] = v2(d.params[2], d.params[3]);
      if (d.flags & 1) {
        s->one_sided = 1;
        s->v[2] = v2(d.params[4], d.params[5]);
        s->v[3] = v2(d.params[6], d.params[7]);
      }
      return B2G_OK;
    case B2G_SHAPE_BOX: {
      host_box(s, d.params[0], d.params[1]);
      if (d.flags & 1) {
        V2 c = v2(d.params[2], d.params[3]);
        s->p = c;
        Xf xf;
        xf.p = c;
        xf.q = rot(d.params[4]);
        for (int i = 0; i < 4; ++i) {
          s->v[i] = mul(xf, s->v[i]);
          s->n[i] = mul(xf.q, s->n[i]);
        }
      }
      return B2G_OK;
    }
    case B2G_SHAPE_POLYGON: {
      s->type = kPolygon;
      s->radius = kPolygonRadius;
      int n = std::min<int>(d.count, 8);
      if (n < 3) { host_box(s, 1.0f, 1.0f); return B2G_ERR_DEGENERATE; }
      // weld near-duplicate points
      V2 ps[8];
      int m = 0;
      const float weld = (0.5f * kLinearSlop) * (0.5f * kLinearSlop);
      for (int i = 0; i < n; ++i) {
        V2 v = v2(d.params[2 * i], d.params[2 * i + 1]);
        bool unique = true;
        for (int j = 0; j < m; ++j)
          if (dist_sq(v, ps[j]) < weld) { unique = false; break; }
        if (unique) ps[m++] = v;
      }
      if (m < 3) { host_box(s, 1.0f, 1.0f); return B2G_ERR_DEGENERATE; }
      // rightmost point (lowest y on ties) starts the wrap
      int i0 = 0;
      float x0 = ps[0].x;
      for (int i = 1; i < m; ++i) {
        float x = ps[i].x;
        if (x > x0 || (x == x0 && ps[i].y < ps[i0].y)) { i0 = i; x0 = x; }
      }
      int hull[8];
      int h = 0, ih = i0;
      for (;;) {
        if (h >= 8) { host_box(s, 1.0f, 1.0f); return B2G_ERR_DEGENERATE; }
        hull[h] = ih;
        int ie = 0;
        for (int j = 1; j < m; ++j) {
          if (ie == ih) { ie = j; continue; }
          V2 r = ps[ie] - ps[hull[h]];
          V2 v = ps[j] - ps[hull[h]];
          float c = cross(r, v);
          if (c < 0.0f) ie = j;
          if (c == 0.0f && dot(v, v) > dot(r, r)) ie = j;
        }
        ++h;
        ih = ie;
        if (ie == i0) break;
      }
      if (h < 3) { host_box(s, 1.0f, 1.0f); return B2G_ERR_DEGENERATE; }
      s->count = h;
      for (int i = 0; i < h; ++i) s->v[i] = ps[hull[i]];
      for (int i = 0; i < h; ++i) {
        int i2 = i + 1 < h ? i + 1 : 0;
        V2 e = s->v[i2] - s->v[i];
        s->n[i] = cross(e, 1.0f);
        normalize(&s->n[i]);
      }
      s->p = host_centroid(s->v, h);
      return B2G_OK;
    }
    case B2G_SHAPE_POLYGON_EXACT: {
      // a stored B2polygonShape taken as it is (the serde polygon record,
      // src/shapes/serialize_b2_polygon_shape.rs:40-90): no hull, normals and centroid given
      const int n = d.count;
      if (n < 3 || n > 8) { host_box(s, 1.0f, 1.0f); return B2G_ERR_DEGENERATE; }
      s->type = kPolygon;
      s->radius = d.radius;
      s->count = n;
      for (int i = 0; i < n; ++i) {
        s->v[i] = v2(d.params[2 * i], d.params[2 * i + 1]);
        s->n[i] = v2(d.params[16 + 2 * i], d.params[16 + 2 * i + 1]);
      }
      s->p = v2(d.params[32], d.params[33]);
      return B2G_OK;
    }
    default:
      return B2G_ERR_UNSUPPORTED;
  }
}

void host_shape_mass(const HostShape& s, float density, float* mass, V2* center, float* I) {
  if (s.type == kCircle) {
    *mass = density * kPi * s.radius * s.radius;
    *center = s.p;
    *I = *mass * (0.5f * s.radius * s.radius + dot(s.p, s.p));
  } else if (s.type == kPolygon) {
    V2 c = v2(0.0f, 0.0f);
    float area = 0.0f, inert = 0.0f;
    const V2 ref = s.v[0];
    const float inv3 = 1.0f / 3.0f;
    for (int i = 0; i < s.count; ++i) {
      V2 e1 = s.v[i] - ref;
      V2 e2 = (i + 1 < s.count) ? s.v[i + 1] - ref : s.v[0] - ref;
      float d = cross(e1, e2);
      float tri = 0.5f * d;
      area += tri;
      c = c + (tri * inv3) * (e1 + e2);
      float intx2 = e1.x * e1.x + e2.x * e1.x + e2.x * e2.x;
      float inty2 = e1.y * e1.y + e2.y * e1.y + e2.y * e2.y;
      inert += ((0.25f * inv3) * d) * (intx2 + inty2);
    }
    *mass = density * area;
    c = (1.0f / area) * c;
    *center = c + ref;
    *I = density * inert;
    *I += *mass * (dot(*center, *center) - dot(c, c));
  } else {
    *mass = 0.0f;
    *center = 0.5f * (s.v[0] + s.v[1]);
    *I = 0.0f;
  }
}

ShapeV host_shape_view(const HostShape& h) {
  ShapeV s;
  s.type = h.type;
  s.count = h.count;
  s.one_sided = h.one_sided;
  s.radius = h.radius;
  s.p = h.p;
  for (int i = 0; i < 8; ++i) { s.v[i] = h.v[i]; s.n[i] = h.n[i]; }
  return s;
}

}  // namespace b2g
