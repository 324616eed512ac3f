// Host build of the GPU narrowphase code (box2d-rs_amd/csrc/b2g_collide.h compiled as plain
// C++), so its logic can be compared with the oracle on the CPU. Test infrastructure only.
// Records use the oracle's float layouts (oracle/b2o_capi.cpp: B2oShape, B2oManifold).
#include <cstring>

#include "../../box2d-rs_amd/csrc/b2g_collide.h"

using namespace b2g;

extern "C" {

struct Shape {
  int32_t type, count, one_sided, pad;
  float radius;
  float p[2], centroid[2], v0[2], v1[2], v2[2], v3[2];
  float vertices[16], normals[16];
};
struct Point {
  float local_point[2];
  float normal_impulse, tangent_impulse;
  uint32_t id;
  int32_t pad;
};
struct Mf {
  int32_t type, point_count;
  float local_normal[2], local_point[2];
  Point points[2];
};

static ShapeV view(const Shape& r) {
  ShapeV s;
  std::memset(&s, 0, sizeof(s));
  s.type = r.type;
  s.count = r.count;
  s.one_sided = r.one_sided;
  s.radius = r.radius;
  if (r.type == kCircle) {
    s.p = v2(r.p[0], r.p[1]);
  } else if (r.type == kPolygon) {
    s.p = v2(r.centroid[0], r.centroid[1]);
    for (int i = 0; i < 8; ++i) {
      s.v[i] = v2(r.vertices[2 * i], r.vertices[2 * i + 1]);
      s.n[i] = v2(r.normals[2 * i], r.normals[2 * i + 1]);
    }
  } else {
    s.v[0] = v2(r.v1[0], r.v1[1]);
    s.v[1] = v2(r.v2[0], r.v2[1]);
    s.v[2] = v2(r.v0[0], r.v0[1]);
    s.v[3] = v2(r.v3[0], r.v3[1]);
  }
  return s;
}

void hc_collide(int n, const Shape* a, const float* xa, const Shape* b, const float* xb, Mf* out) {
  for (int i = 0; i < n; ++i) {
    Xf ta, tb;
    ta.p = v2(xa[4 * i], xa[4 * i + 1]); ta.q.s = xa[4 * i + 2]; ta.q.c = xa[4 * i + 3];
    tb.p = v2(xb[4 * i], xb[4 * i + 1]); tb.q.s = xb[4 * i + 2]; tb.q.c = xb[4 * i + 3];
    Manifold m;
    evaluate(&m, view(a[i]), ta, view(b[i]), tb);
    Mf& o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.type = m.type;
    o.point_count = m.count;
    o.local_normal[0] = m.local_normal.x; o.local_normal[1] = m.local_normal.y;
    o.local_point[0] = m.local_point.x; o.local_point[1] = m.local_point.y;
    for (int p = 0; p < 2; ++p) {
      o.points[p].local_point[0] = m.lp[p].x;
      o.points[p].local_point[1] = m.lp[p].y;
      o.points[p].id = m.id[p];
    }
  }
}

void hc_aabb(int n, const Shape* s, const float* x, float* out) {
  for (int i = 0; i < n; ++i) {
    Xf t;
    t.p = v2(x[4 * i], x[4 * i + 1]); t.q.s = x[4 * i + 2]; t.q.c = x[4 * i + 3];
    Aabb a = shape_aabb(view(s[i]), t);
    out[4 * i] = a.lo.x; out[4 * i + 1] = a.lo.y; out[4 * i + 2] = a.hi.x; out[4 * i + 3] = a.hi.y;
  }
}

}  // extern "C"
