// GJK distance between two convex shapes (src/private/collision/b2_distance.rs:12-41 proxies,
// 404-530 b2_distance_fn), shared by the time of impact and the sensors' overlap test
// (src/private/collision/b2_collision.rs:216-245).
#pragma once
#include "b2g_collide.h"

namespace b2g {

// A shape as GJK sees it: its core points and radius.
struct Proxy {
  const V2* v;
  int count;
  float radius;
  __device__ int support(V2 d) const {
    int best = 0;
    float best_value = dot(v[0], d);
    for (int i = 1; i < count; ++i) {
      const float value = dot(v[i], d);
      if (value > best_value) {
        best = i;
        best_value = value;
      }
    }
    return best;
  }
};

__device__ inline Proxy proxy_of(const ShapeV& s) {
  Proxy p;
  if (s.type == kCircle) {
    p.v = &s.p;
    p.count = 1;
  } else if (s.type == kPolygon) {
    p.v = s.v;
    p.count = s.count;
  } else {
    p.v = s.v;  // vertex1, vertex2
    p.count = 2;
  }
  p.radius = s.radius;
  return p;
}

struct Cache {
  float metric;
  int count;
  int ia[3], ib[3];
};

struct SVert {
  V2 wa, wb, w;
  float a;
  int ia, ib;
};

struct Simplex {
  SVert v[3];
  int count;

  __device__ float metric() const {
    if (count == 2) return len(v[0].w - v[1].w);
    if (count == 3) return cross(v[1].w - v[0].w, v[2].w - v[0].w);
    return 0.0f;
  }
  __device__ void vertex(int i, int ia, int ib, const Proxy& pa, const Xf& xa, const Proxy& pb, const Xf& xb) {
    v[i].ia = ia;
    v[i].ib = ib;
    v[i].wa = mul(xa, pa.v[ia]);
    v[i].wb = mul(xb, pb.v[ib]);
    v[i].w = v[i].wb - v[i].wa;
  }
  __device__ void read_cache(const Cache& c, const Proxy& pa, const Xf& xa, const Proxy& pb, const Xf& xb) {
    count = c.count;
    for (int i = 0; i < count; ++i) {
      vertex(i, c.ia[i], c.ib[i], pa, xa, pb, xb);
      v[i].a = 0.0f;
    }
    if (count > 1) {
      const float m1 = c.metric, m2 = metric();
      if (m2 < 0.5f * m1 || 2.0f * m1 < m2 || m2 < kEpsilon) count = 0;
    }
    if (count == 0) {
      vertex(0, 0, 0, pa, xa, pb, xb);
      v[0].a = 1.0f;
      count = 1;
    }
  }
  __device__ void write_cache(Cache& c) const {
    c.metric = metric();
    c.count = count;
    for (int i = 0; i < count; ++i) {
      c.ia[i] = v[i].ia;
      c.ib[i] = v[i].ib;
    }
  }
  __device__ void solve2() {
    const V2 w1 = v[0].w, w2 = v[1].w, e12 = w2 - w1;
    const float d12_2 = -dot(w1, e12);
    if (d12_2 <= 0.0f) {
      v[0].a = 1.0f;
      count = 1;
      return;
    }
    const float d12_1 = dot(w2, e12);
    if (d12_1 <= 0.0f) {
      v[1].a = 1.0f;
      count = 1;
      v[0] = v[1];
      return;
    }
    const float inv = 1.0f / (d12_1 + d12_2);
    v[0].a = d12_1 * inv;
    v[1].a = d12_2 * inv;
    count = 2;
  }
  __device__ void solve3() {
    const V2 w1 = v[0].w, w2 = v[1].w, w3 = v[2].w;
    const V2 e12 = w2 - w1, e13 = w3 - w1, e23 = w3 - w2;
    const float d12_1 = dot(w2, e12), d12_2 = -dot(w1, e12);
    const float d13_1 = dot(w3, e13), d13_2 = -dot(w1, e13);
    const float d23_1 = dot(w3, e23), d23_2 = -dot(w2, e23);
    const float n123 = cross(e12, e13);
    const float d123_1 = n123 * cross(w2, w3), d123_2 = n123 * cross(w3, w1), d123_3 = n123 * cross(w1, w2);
    if (d12_2 <= 0.0f && d13_2 <= 0.0f) {
      v[0].a = 1.0f;
      count = 1;
    } else if (d12_1 > 0.0f && d12_2 > 0.0f && d123_3 <= 0.0f) {
      const float inv = 1.0f / (d12_1 + d12_2);
      v[0].a = d12_1 * inv;
      v[1].a = d12_2 * inv;
      count = 2;
    } else if (d13_1 > 0.0f && d13_2 > 0.0f && d123_2 <= 0.0f) {
      const float inv = 1.0f / (d13_1 + d13_2);
      v[0].a = d13_1 * inv;
      v[2].a = d13_2 * inv;
      count = 2;
      v[1] = v[2];
    } else if (d12_1 <= 0.0f && d23_2 <= 0.0f) {
      v[1].a = 1.0f;
      count = 1;
      v[0] = v[1];
    } else if (d13_1 <= 0.0f && d23_1 <= 0.0f) {
      v[2].a = 1.0f;
      count = 1;
      v[0] = v[2];
    } else if (d23_1 > 0.0f && d23_2 > 0.0f && d123_1 <= 0.0f) {
      const float inv = 1.0f / (d23_1 + d23_2);
      v[1].a = d23_1 * inv;
      v[2].a = d23_2 * inv;
      count = 2;
      v[0] = v[2];
    } else {
      const float inv = 1.0f / (d123_1 + d123_2 + d123_3);
      v[0].a = d123_1 * inv;
      v[1].a = d123_2 * inv;
      v[2].a = d123_3 * inv;
      count = 3;
    }
  }
};

// GJK distance between the core shapes (no radii), with a simplex cache carried between calls.
__device__ inline float gjk_distance(Cache& cache, const Proxy& pa, const Xf& xa, const Proxy& pb, const Xf& xb) {
  Simplex s;
  s.read_cache(cache, pa, xa, pb, xb);
  int save_a[3], save_b[3];
  for (int iter = 0; iter < 20;) {
    const int save_count = s.count;
    for (int i = 0; i < save_count; ++i) {
      save_a[i] = s.v[i].ia;
      save_b[i] = s.v[i].ib;
    }
    if (s.count == 2) s.solve2();
    else if (s.count == 3) s.solve3();
    if (s.count == 3) break;
    V2 d;
    if (s.count == 1) {
      d = -s.v[0].w;
    } else {
      const V2 e12 = s.v[1].w - s.v[0].w;
      d = cross(e12, -s.v[0].w) > 0.0f ? cross(1.0f, e12) : cross(e12, 1.0f);
    }
    if (dot(d, d) < kEpsilon * kEpsilon) break;
    const int ia = pa.support(mulT(xa.q, -d));
    const int ib = pb.support(mulT(xb.q, d));
    s.vertex(s.count, ia, ib, pa, xa, pb, xb);
    ++iter;
    bool duplicate = false;
    for (int i = 0; i < save_count; ++i)
      if (ia == save_a[i] && ib == save_b[i]) duplicate = true;
    if (duplicate) break;
    ++s.count;
  }
  V2 wa, wb;
  if (s.count == 1) {
    wa = s.v[0].wa;
    wb = s.v[0].wb;
  } else if (s.count == 2) {
    wa = s.v[0].a * s.v[0].wa + s.v[1].a * s.v[1].wa;
    wb = s.v[0].a * s.v[0].wb + s.v[1].a * s.v[1].wb;
  } else {
    wa = s.v[0].a * s.v[0].wa + s.v[1].a * s.v[1].wa + s.v[2].a * s.v[2].wa;
    wb = wa;
  }
  s.write_cache(cache);
  return len(wa - wb);
}

// b2_test_overlap: the distance with radii from an empty simplex cache is below ten epsilons
// (b2_distance_fn's use_radii: the core distance less both radii when it exceeds them, else 0)
__device__ inline bool test_overlap(const ShapeV& a, const Xf& xa, const ShapeV& b, const Xf& xb) {
  Cache cache;
  cache.metric = 0.0f;
  cache.count = 0;
  const Proxy pa = proxy_of(a), pb = proxy_of(b);
  float d = gjk_distance(cache, pa, xa, pb, xb);
  const float rab = pa.radius + pb.radius;
  if (d > rab && d > kEpsilon) d -= rab;
  else d = 0.0f;
  return d < 10.0f * kEpsilon;
}

}  // namespace b2g
