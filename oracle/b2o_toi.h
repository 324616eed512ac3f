// Oracle (test infrastructure only): GJK distance and conservative-advancement time of impact,
// restated from the reference for the continuous-collision checks.
//   sweep: src/b2_math.rs:764-787 (get_transform, advance, normalize)
//   distance proxy: src/private/collision/b2_distance.rs:12-41, support src/b2_distance.rs:143-155
//   GJK: src/private/collision/b2_distance.rs:60-530 (simplex, solve2, solve3, b2_distance_fn)
//   separation function and TOI: src/private/collision/b2_time_of_impact.rs:17-463
#pragma once
#include "b2o_core.h"

namespace b2o {

// src/b2_math.rs:773-779
inline void sweep_advance(Sweep* s, real alpha) {
  real beta = (alpha - s->alpha0) / (real(1.0) - s->alpha0);
  s->c0 += beta * (s->c - s->c0);
  s->a0 += beta * (s->a - s->a0);
  s->alpha0 = alpha;
}

// src/b2_math.rs:782-787
inline void sweep_normalize(Sweep* s) {
  real two_pi = real(2.0) * kPi;
  real d = two_pi * std::floor(s->a0 / two_pi);
  s->a0 -= d;
  s->a -= d;
}

// src/private/collision/b2_distance.rs:12-41 (set_shape; chains are not part of this path)
struct DistanceProxy {
  Vec2 v[kMaxPolygonVertices];
  int count = 0;
  real radius = 0;
  void set(const Shape& s) {
    if (s.type == kCircle) {
      v[0] = s.p;
      count = 1;
    } else if (s.type == kPolygon) {
      for (int i = 0; i < s.count; ++i) v[i] = s.vertices[i];
      count = s.count;
    } else {
      v[0] = s.v1;
      v[1] = s.v2;
      count = 2;
    }
    radius = s.radius;
  }
  // src/b2_distance.rs:143-155: the first vertex of the largest projection
  int support(Vec2 d) const {
    int best = 0;
    real best_value = dot(v[0], d);
    for (int i = 1; i < count; ++i) {
      real value = dot(v[i], d);
      if (value > best_value) {
        best = i;
        best_value = value;
      }
    }
    return best;
  }
};

struct SimplexCache {
  real metric = 0;
  int count = 0;
  int index_a[3] = {0, 0, 0}, index_b[3] = {0, 0, 0};
};

struct SimplexVertex {
  Vec2 wa, wb, w;
  real a = 0;
  int ia = 0, ib = 0;
};

struct Simplex {
  SimplexVertex v[3];
  int count = 0;

  real metric() const {
    if (count == 2) return std::sqrt((v[0].w - v[1].w).length_squared());
    if (count == 3) return cross(v[1].w - v[0].w, v[2].w - v[0].w);
    return 0;
  }

  // src/private/collision/b2_distance.rs:71-120
  void read_cache(const SimplexCache& cache, const DistanceProxy& pa, const Transform& xa, const DistanceProxy& pb,
                  const Transform& xb) {
    count = cache.count;
    for (int i = 0; i < count; ++i) {
      SimplexVertex& s = v[i];
      s.ia = cache.index_a[i];
      s.ib = cache.index_b[i];
      s.wa = mul(xa, pa.v[s.ia]);
      s.wb = mul(xb, pb.v[s.ib]);
      s.w = s.wb - s.wa;
      s.a = 0;
    }
    if (count > 1) {
      real metric1 = cache.metric, metric2 = metric();
      if (metric2 < real(0.5) * metric1 || real(2.0) * metric1 < metric2 || metric2 < kEpsilon) count = 0;
    }
    if (count == 0) {
      SimplexVertex& s = v[0];
      s.ia = 0;
      s.ib = 0;
      s.wa = mul(xa, pa.v[0]);
      s.wb = mul(xb, pb.v[0]);
      s.w = s.wb - s.wa;
      s.a = 1;
      count = 1;
    }
  }

  void write_cache(SimplexCache* cache) const {
    cache->metric = metric();
    cache->count = count;
    for (int i = 0; i < count; ++i) {
      cache->index_a[i] = v[i].ia;
      cache->index_b[i] = v[i].ib;
    }
  }

  // src/private/collision/b2_distance.rs:134-156
  Vec2 search_direction() const {
    if (count == 1) return -v[0].w;
    Vec2 e12 = v[1].w - v[0].w;
    real sgn = cross(e12, -v[0].w);
    return sgn > 0 ? cross(real(1.0), e12) : cross(e12, real(1.0));
  }

  // src/private/collision/b2_distance.rs:182-210
  void witness_points(Vec2* pa, Vec2* pb) const {
    if (count == 1) {
      *pa = v[0].wa;
      *pb = v[0].wb;
    } else if (count == 2) {
      *pa = v[0].a * v[0].wa + v[1].a * v[1].wa;
      *pb = v[0].a * v[0].wb + v[1].a * v[1].wb;
    } else {
      *pa = v[0].a * v[0].wa + v[1].a * v[1].wa + v[2].a * v[2].wa;
      *pb = *pa;
    }
  }

  // src/private/collision/b2_distance.rs:264-293
  void solve2() {
    Vec2 w1 = v[0].w, w2 = v[1].w;
    Vec2 e12 = w2 - w1;
    real d12_2 = -dot(w1, e12);
    if (d12_2 <= 0) {
      v[0].a = 1;
      count = 1;
      return;
    }
    real d12_1 = dot(w2, e12);
    if (d12_1 <= 0) {
      v[1].a = 1;
      count = 1;
      v[0] = v[1];
      return;
    }
    real inv = real(1.0) / (d12_1 + d12_2);
    v[0].a = d12_1 * inv;
    v[1].a = d12_2 * inv;
    count = 2;
  }

  // src/private/collision/b2_distance.rs:300-402
  void solve3() {
    Vec2 w1 = v[0].w, w2 = v[1].w, w3 = v[2].w;
    Vec2 e12 = w2 - w1;
    real d12_1 = dot(w2, e12), d12_2 = -dot(w1, e12);
    Vec2 e13 = w3 - w1;
    real d13_1 = dot(w3, e13), d13_2 = -dot(w1, e13);
    Vec2 e23 = w3 - w2;
    real d23_1 = dot(w3, e23), d23_2 = -dot(w2, e23);
    real n123 = cross(e12, e13);
    real d123_1 = n123 * cross(w2, w3);
    real d123_2 = n123 * cross(w3, w1);
    real d123_3 = n123 * cross(w1, w2);
    if (d12_2 <= 0 && d13_2 <= 0) {
      v[0].a = 1;
      count = 1;
      return;
    }
    if (d12_1 > 0 && d12_2 > 0 && d123_3 <= 0) {
      real inv = real(1.0) / (d12_1 + d12_2);
      v[0].a = d12_1 * inv;
      v[1].a = d12_2 * inv;
      count = 2;
      return;
    }
    if (d13_1 > 0 && d13_2 > 0 && d123_2 <= 0) {
      real inv = real(1.0) / (d13_1 + d13_2);
      v[0].a = d13_1 * inv;
      v[2].a = d13_2 * inv;
      count = 2;
      v[1] = v[2];
      return;
    }
    if (d12_1 <= 0 && d23_2 <= 0) {
      v[1].a = 1;
      count = 1;
      v[0] = v[1];
      return;
    }
    if (d13_1 <= 0 && d23_1 <= 0) {
      v[2].a = 1;
      count = 1;
      v[0] = v[2];
      return;
    }
    if (d23_1 > 0 && d23_2 > 0 && d123_1 <= 0) {
      real inv = real(1.0) / (d23_1 + d23_2);
      v[1].a = d23_1 * inv;
      v[2].a = d23_2 * inv;
      count = 2;
      v[0] = v[2];
      return;
    }
    real inv = real(1.0) / (d123_1 + d123_2 + d123_3);
    v[0].a = d123_1 * inv;
    v[1].a = d123_2 * inv;
    v[2].a = d123_3 * inv;
    count = 3;
  }
};

struct DistanceOutput {
  Vec2 pa, pb;
  real distance = 0;
  int iterations = 0;
};

// src/private/collision/b2_distance.rs:404-530 (b2_distance_fn)
inline void distance(DistanceOutput* out, SimplexCache* cache, const DistanceProxy& pa, const Transform& xa,
                     const DistanceProxy& pb, const Transform& xb, bool use_radii) {
  Simplex s;
  s.read_cache(*cache, pa, xa, pb, xb);
  const int kMaxIters = 20;
  int save_a[3], save_b[3];
  int iter = 0;
  while (iter < kMaxIters) {
    int save_count = s.count;
    for (int i = 0; i < save_count; ++i) {
      save_a[i] = s.v[i].ia;
      save_b[i] = s.v[i].ib;
    }
    if (s.count == 2) s.solve2();
    else if (s.count == 3) s.solve3();
    if (s.count == 3) break;
    Vec2 d = s.search_direction();
    if (d.length_squared() < kEpsilon * kEpsilon) break;
    SimplexVertex& vx = s.v[s.count];
    vx.ia = pa.support(mulT(xa.q, -d));
    vx.wa = mul(xa, pa.v[vx.ia]);
    vx.ib = pb.support(mulT(xb.q, d));
    vx.wb = mul(xb, pb.v[vx.ib]);
    vx.w = vx.wb - vx.wa;
    ++iter;
    bool duplicate = false;
    for (int i = 0; i < save_count; ++i) {
      if (vx.ia == save_a[i] && vx.ib == save_b[i]) {
        duplicate = true;
        break;
      }
    }
    if (duplicate) break;
    ++s.count;
  }
  s.witness_points(&out->pa, &out->pb);
  out->distance = std::sqrt((out->pa - out->pb).length_squared());
  out->iterations = iter;
  s.write_cache(cache);
  if (use_radii) {
    real ra = pa.radius, rb = pb.radius;
    if (out->distance > ra + rb && out->distance > kEpsilon) {
      out->distance -= ra + rb;
      Vec2 n = out->pb - out->pa;
      n.normalize();
      out->pa += ra * n;
      out->pb -= rb * n;
    } else {
      Vec2 p = real(0.5) * (out->pa + out->pb);
      out->pa = p;
      out->pb = p;
      out->distance = 0;
    }
  }
}

// src/private/collision/b2_time_of_impact.rs:17-230
struct SeparationFunction {
  enum { kPoints = 0, kFaceA = 1, kFaceB = 2 };
  const DistanceProxy *pa, *pb;
  Sweep sa, sb;
  int type = kPoints;
  Vec2 local_point, axis;

  void initialize(const SimplexCache& cache, const DistanceProxy* proxy_a, const Sweep& sweep_a,
                  const DistanceProxy* proxy_b, const Sweep& sweep_b, real t1) {
    pa = proxy_a;
    pb = proxy_b;
    sa = sweep_a;
    sb = sweep_b;
    Transform xa, xb;
    sa.get_transform(&xa, t1);
    sb.get_transform(&xb, t1);
    if (cache.count == 1) {
      type = kPoints;
      Vec2 qa = mul(xa, pa->v[cache.index_a[0]]);
      Vec2 qb = mul(xb, pb->v[cache.index_b[0]]);
      axis = qb - qa;
      axis.normalize();
    } else if (cache.index_a[0] == cache.index_a[1]) {
      type = kFaceB;
      Vec2 b1 = pb->v[cache.index_b[0]], b2 = pb->v[cache.index_b[1]];
      axis = cross(b2 - b1, real(1.0));
      axis.normalize();
      Vec2 normal = mul(xb.q, axis);
      local_point = real(0.5) * (b1 + b2);
      Vec2 qb = mul(xb, local_point);
      Vec2 qa = mul(xa, pa->v[cache.index_a[0]]);
      real s = dot(qa - qb, normal);
      if (s < 0) axis = -axis;
    } else {
      type = kFaceA;
      Vec2 a1 = pa->v[cache.index_a[0]], a2 = pa->v[cache.index_a[1]];
      axis = cross(a2 - a1, real(1.0));
      axis.normalize();
      Vec2 normal = mul(xa.q, axis);
      local_point = real(0.5) * (a1 + a2);
      Vec2 qa = mul(xa, local_point);
      Vec2 qb = mul(xb, pb->v[cache.index_b[0]]);
      real s = dot(qb - qa, normal);
      if (s < 0) axis = -axis;
    }
  }

  real find_min_separation(int* ia, int* ib, real t) const {
    Transform xa, xb;
    sa.get_transform(&xa, t);
    sb.get_transform(&xb, t);
    if (type == kPoints) {
      *ia = pa->support(mulT(xa.q, axis));
      *ib = pb->support(mulT(xb.q, -axis));
      Vec2 qa = mul(xa, pa->v[*ia]);
      Vec2 qb = mul(xb, pb->v[*ib]);
      return dot(qb - qa, axis);
    }
    if (type == kFaceA) {
      Vec2 normal = mul(xa.q, axis);
      Vec2 qa = mul(xa, local_point);
      *ia = -1;
      *ib = pb->support(mulT(xb.q, -normal));
      Vec2 qb = mul(xb, pb->v[*ib]);
      return dot(qb - qa, normal);
    }
    Vec2 normal = mul(xb.q, axis);
    Vec2 qb = mul(xb, local_point);
    *ib = -1;
    *ia = pa->support(mulT(xa.q, -normal));
    Vec2 qa = mul(xa, pa->v[*ia]);
    return dot(qa - qb, normal);
  }

  real evaluate(int ia, int ib, real t) const {
    Transform xa, xb;
    sa.get_transform(&xa, t);
    sb.get_transform(&xb, t);
    if (type == kPoints) {
      Vec2 qa = mul(xa, pa->v[ia]);
      Vec2 qb = mul(xb, pb->v[ib]);
      return dot(qb - qa, axis);
    }
    if (type == kFaceA) {
      Vec2 normal = mul(xa.q, axis);
      Vec2 qa = mul(xa, local_point);
      Vec2 qb = mul(xb, pb->v[ib]);
      return dot(qb - qa, normal);
    }
    Vec2 normal = mul(xb.q, axis);
    Vec2 qb = mul(xb, local_point);
    Vec2 qa = mul(xa, pa->v[ia]);
    return dot(qa - qb, normal);
  }
};

enum ToiState { kToiUnknown = 0, kToiFailed = 1, kToiOverlapped = 2, kToiTouching = 3, kToiSeparated = 4 };
struct ToiOutput {
  int state = kToiUnknown;
  real t = 0;
};

// src/private/collision/b2_time_of_impact.rs:232-463
inline void time_of_impact(ToiOutput* out, const DistanceProxy& pa, const DistanceProxy& pb, Sweep sweep_a,
                           Sweep sweep_b, real t_max) {
  out->state = kToiUnknown;
  out->t = t_max;
  sweep_normalize(&sweep_a);
  sweep_normalize(&sweep_b);
  real total_radius = pa.radius + pb.radius;
  real target = bmax(kLinearSlop, total_radius - real(3.0) * kLinearSlop);
  real tolerance = real(0.25) * kLinearSlop;
  real t1 = 0;
  const int kMaxIterations = 20;
  int iter = 0;
  SimplexCache cache;
  for (;;) {
    Transform xa, xb;
    sweep_a.get_transform(&xa, t1);
    sweep_b.get_transform(&xb, t1);
    DistanceOutput d;
    distance(&d, &cache, pa, xa, pb, xb, false);
    if (d.distance <= 0) {
      out->state = kToiOverlapped;
      out->t = 0;
      break;
    }
    if (d.distance < target + tolerance) {
      out->state = kToiTouching;
      out->t = t1;
      break;
    }
    SeparationFunction fcn;
    fcn.initialize(cache, &pa, sweep_a, &pb, sweep_b, t1);
    bool done = false;
    real t2 = t_max;
    int push_back_iter = 0;
    for (;;) {
      int ia = 0, ib = 0;
      real s2 = fcn.find_min_separation(&ia, &ib, t2);
      if (s2 > target + tolerance) {
        out->state = kToiSeparated;
        out->t = t_max;
        done = true;
        break;
      }
      if (s2 > target - tolerance) {
        t1 = t2;
        break;
      }
      real s1 = fcn.evaluate(ia, ib, t1);
      if (s1 < target - tolerance) {
        out->state = kToiFailed;
        out->t = t1;
        done = true;
        break;
      }
      if (s1 <= target + tolerance) {
        out->state = kToiTouching;
        out->t = t1;
        done = true;
        break;
      }
      int root_iters = 0;
      real a1 = t1, a2 = t2;
      for (;;) {
        real t;
        if (root_iters & 1) t = a1 + (target - s1) * (a2 - a1) / (s2 - s1);
        else t = real(0.5) * (a1 + a2);
        ++root_iters;
        real s = fcn.evaluate(ia, ib, t);
        if (babs(s - target) < tolerance) {
          t2 = t;
          break;
        }
        if (s > target) {
          a1 = t;
          s1 = s;
        } else {
          a2 = t;
          s2 = s;
        }
        if (root_iters == 50) break;
      }
      ++push_back_iter;
      if (push_back_iter == kMaxPolygonVertices) break;
    }
    ++iter;
    if (done) break;
    if (iter == kMaxIterations) {
      out->state = kToiFailed;
      out->t = t1;
      break;
    }
  }
}

}  // namespace b2o
