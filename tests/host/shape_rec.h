// Shape records in the oracle's float layout (oracle/b2o_capi.cpp: B2oShape) and their view
// as the library's ShapeV, for the host checks. Test infrastructure only.
#pragma once
#include <cstring>

#include "../../box2d-rs_amd/csrc/b2g_collide.h"

namespace shape_rec {
using namespace b2g;

struct Shape {
  int32_t type, count, one_sided, pad;
  float radius;
  float p[2], centroid[2], v0[2], v1[2], v2[2], v3[2];
  float vertices[16], normals[16];
};
inline ShapeV view(const Shape& r) {
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

}  // namespace shape_rec
