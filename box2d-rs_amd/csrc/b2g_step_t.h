// b2g: the per-body and per-constraint functions of the island solve, templated on the scalar.
//
// The same operations, in the same order, as the simulation kernels of b2g_solver.hip (which
// are specialised for the SoA float layout); instantiated with float they reproduce those
// kernels, with double they give the host's f64 model, and with Dual (b2g_dual.h) the local
// Jacobians of the adjoint. References: src/private/dynamics/b2_island_private.rs:135-252 and
// src/private/dynamics/b2_contact_solver_private.rs:113-730.
#pragma once
#include "b2g_collide.h"

namespace b2g {

B2G_HD float tabs(float x) { return fabsf(x); }
B2G_HD double tabs(double x) { return fabs(x); }
template <class T>
B2G_HD T tabs(const T& x) { return x < T(0.0) ? -x : x; }

// constant per-body data of the solver (S: float on the device, double in host models)
template <class S>
struct TBodyMass {
  S inv_m, inv_i, lcx, lcy;  // inverse mass, inverse inertia, local centre
  S mass, gravity_scale, lin_damping, ang_damping;
};
using BodyMass = TBodyMass<float>;

// position (sweep c, a) and velocity (v, w) of one body
template <class T>
struct TPos {
  TV2<T> c;
  T a;
};
template <class T>
struct TVel {
  TV2<T> v;
  T w;
};

// constant per-contact data: manifold type and point count, mixing, radii, inverse masses
template <class S>
struct TContactConst {
  int type, count;
  S restitution, threshold;  // mixed (friction is an argument: it is differentiated)
  S rA, rB;                  // shape radii
  S mA, iA, mB, iB;          // inverse masses
  S lcA[2], lcB[2];          // local centres
};
using ContactConst = TContactConst<float>;

// manifold geometry in the bodies' frames (the part that depends on the transforms)
template <class T>
struct TGeom {
  TV2<T> ln, lpt, lp[2];
};

// B2contactVelocityConstraint as initialize_velocity_constraints leaves it
template <class T>
struct TVelC {
  TV2<T> normal;
  T friction;
  TV2<T> ra[2], rb[2];
  T nm[2], tm[2], bias[2];
  T K[4], NM[4];  // ex.x, ex.y, ey.x, ey.y
  int vcount;
};

// src/private/dynamics/b2_island_private.rs:135-170 (dynamic bodies)
template <class T, class S>
B2G_HD void integrate_velocity_t(TVel<T>& v, const TBodyMass<S>& m, S fx, S fy, S torque, S gx, S gy, S h) {
  T s = T(h) * T(m.inv_m);
  T gm = T(m.gravity_scale) * T(m.mass);
  v.v.x += s * (gm * T(gx) + T(fx));
  v.v.y += s * (gm * T(gy) + T(fy));
  v.w += T(h) * T(m.inv_i) * T(torque);
  T dl = T(1.0) / (T(1.0) + T(h) * T(m.lin_damping));
  T da = T(1.0) / (T(1.0) + T(h) * T(m.ang_damping));
  v.v.x *= dl;
  v.v.y *= dl;
  v.w *= da;
}

// src/private/dynamics/b2_island_private.rs:221-252 (non-static bodies)
template <class T, class S>
B2G_HD void integrate_position_t(TPos<T>& p, TVel<T>& v, S h) {
  TV2<T> tr = T(h) * v.v;
  if (dot(tr, tr) > T(kMaxTranslationD) * T(kMaxTranslationD)) {
    T ratio = T(kMaxTranslationD) / len(tr);
    v.v.x *= ratio;
    v.v.y *= ratio;
  }
  T rotation = T(h) * v.w;
  const T maxRot = T(0.5 * kPiD);
  if (rotation * rotation > maxRot * maxRot) {
    T ratio = maxRot / tabs(rotation);
    v.w *= ratio;
  }
  p.c.x += T(h) * v.v.x;
  p.c.y += T(h) * v.v.y;
  p.a += T(h) * v.w;
}

// body transform from the sweep (synchronize_transform, src/b2_body.rs:974-977)
template <class T, class S>
B2G_HD TXf<T> xf_of_t(const TPos<T>& p, const S lc[2]) {
  TXf<T> xf;
  xf.q = rot(p.a);
  xf.p = p.c - mul(xf.q, v2(T(lc[0]), T(lc[1])));
  return xf;
}

// initialize_velocity_constraints for one contact
// (src/private/dynamics/b2_contact_solver_private.rs:113-226)
template <class T, class S>
B2G_HD TVelC<T> init_velocity_t(const TContactConst<S>& k, const TGeom<T>& g, T friction, const TPos<T>& pA,
                                const TPos<T>& pB, const TVel<T>& vA, const TVel<T>& vB, bool block_solve) {
  TVelC<T> c;
  c.friction = friction;
  TXf<T> xfA = xf_of_t(pA, k.lcA), xfB = xf_of_t(pB, k.lcB);
  TWorldManifold<T> wm;
  world_manifold(&wm, k.type, k.count, g.ln, g.lpt, g.lp, xfA, T(k.rA), xfB, T(k.rB));
  c.normal = wm.normal;
  const T mA = T(k.mA), iA = T(k.iA), mB = T(k.mB), iB = T(k.iB);
  for (int p = 0; p < 2; ++p) {
    c.ra[p] = c.rb[p] = v2(T(0.0), T(0.0));
    c.nm[p] = c.tm[p] = c.bias[p] = T(0.0);
  }
  for (int p = 0; p < k.count; ++p) {
    TV2<T> ra = wm.points[p] - pA.c;
    TV2<T> rb = wm.points[p] - pB.c;
    c.ra[p] = ra;
    c.rb[p] = rb;
    T rnA = cross(ra, c.normal), rnB = cross(rb, c.normal);
    T kNormal = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
    c.nm[p] = kNormal > T(0.0) ? T(1.0) / kNormal : T(0.0);
    TV2<T> tangent = cross(c.normal, T(1.0));
    T rtA = cross(ra, tangent), rtB = cross(rb, tangent);
    T kTangent = mA + mB + iA * rtA * rtA + iB * rtB * rtB;
    c.tm[p] = kTangent > T(0.0) ? T(1.0) / kTangent : T(0.0);
    T vRel = dot(c.normal, ((vB.v + cross(vB.w, rb)) - vA.v) - cross(vA.w, ra));
    if (vRel < -T(k.threshold)) c.bias[p] = -T(k.restitution) * vRel;
  }
  c.vcount = k.count;
  for (int i = 0; i < 4; ++i) c.K[i] = c.NM[i] = T(0.0);
  if (k.count == 2 && block_solve) {
    TV2<T> r1a = c.ra[0], r1b = c.rb[0], r2a = c.ra[1], r2b = c.rb[1];
    T rn1A = cross(r1a, c.normal), rn1B = cross(r1b, c.normal);
    T rn2A = cross(r2a, c.normal), rn2B = cross(r2b, c.normal);
    T k11 = mA + mB + iA * rn1A * rn1A + iB * rn1B * rn1B;
    T k22 = mA + mB + iA * rn2A * rn2A + iB * rn2B * rn2B;
    T k12 = mA + mB + iA * rn1A * rn2A + iB * rn1B * rn2B;
    if (k11 * k11 < T(1000.0) * (k11 * k22 - k12 * k12)) {
      c.K[0] = k11; c.K[1] = k12; c.K[2] = k12; c.K[3] = k22;
      T det = k11 * k22 - k12 * k12;  // B2Mat22::get_inverse (src/b2_math.rs:261-274)
      if (det != T(0.0)) det = T(1.0) / det;
      c.NM[0] = det * k22; c.NM[1] = -det * k12; c.NM[2] = -det * k12; c.NM[3] = det * k11;
    } else {
      c.vcount = 1;
    }
  }
  return c;
}

// warm_start for one constraint (src/private/dynamics/b2_contact_solver_private.rs:228-266)
template <class T, class S>
B2G_HD void warm_start_t(const TVelC<T>& c, const TContactConst<S>& k, const T lam[4], TVel<T>& A, TVel<T>& B) {
  TV2<T> tangent = cross(c.normal, T(1.0));
  for (int p = 0; p < c.vcount; ++p) {
    T ni = p == 0 ? lam[0] : lam[2], ti = p == 0 ? lam[1] : lam[3];
    TV2<T> P = ni * c.normal + ti * tangent;
    A.w -= T(k.iA) * cross(c.ra[p], P);
    A.v = A.v - T(k.mA) * P;
    B.w += T(k.iB) * cross(c.rb[p], P);
    B.v = B.v + T(k.mB) * P;
  }
}

template <class T, class S>
B2G_HD void apply_normal_pair_t(TVel<T>& A, TVel<T>& B, const TContactConst<S>& k, const TVelC<T>& c, TV2<T> d) {
  TV2<T> P1 = d.x * c.normal, P2 = d.y * c.normal;
  A.v = A.v - T(k.mA) * (P1 + P2);
  A.w -= T(k.iA) * (cross(c.ra[0], P1) + cross(c.ra[1], P2));
  B.v = B.v + T(k.mB) * (P1 + P2);
  B.w += T(k.iB) * (cross(c.rb[0], P1) + cross(c.rb[1], P2));
}

// solve_velocity_constraints for one constraint, friction then normal (single points or the
// 2-point block LCP by total enumeration): src/private/dynamics/b2_contact_solver_private.rs:268-583
template <class T, class S>
B2G_HD void solve_velocity_t(const TVelC<T>& c, const TContactConst<S>& k, T lam[4], TVel<T>& A, TVel<T>& B,
                             bool block_solve) {
  const T mA = T(k.mA), iA = T(k.iA), mB = T(k.mB), iB = T(k.iB);
  TV2<T> tangent = cross(c.normal, T(1.0));
  const T tangent_speed = T(0.0);
  for (int p = 0; p < c.vcount; ++p) {
    TV2<T> ra = c.ra[p], rb = c.rb[p];
    TV2<T> dv = ((B.v + cross(B.w, rb)) - A.v) - cross(A.w, ra);
    T vt = dot(dv, tangent) - tangent_speed;
    T lambda = c.tm[p] * (-vt);
    T maxFriction = c.friction * lam[2 * p];
    T newImpulse = bclamp(lam[2 * p + 1] + lambda, -maxFriction, maxFriction);
    lambda = newImpulse - lam[2 * p + 1];
    lam[2 * p + 1] = newImpulse;
    TV2<T> P = lambda * tangent;
    A.v = A.v - mA * P;
    A.w -= iA * cross(ra, P);
    B.v = B.v + mB * P;
    B.w += iB * cross(rb, P);
  }
  if (c.vcount == 1 || !block_solve) {
    for (int p = 0; p < c.vcount; ++p) {
      TV2<T> ra = c.ra[p], rb = c.rb[p];
      TV2<T> dv = ((B.v + cross(B.w, rb)) - A.v) - cross(A.w, ra);
      T vn = dot(dv, c.normal);
      T lambda = -c.nm[p] * (vn - c.bias[p]);
      T newImpulse = bmax(lam[2 * p] + lambda, T(0.0));
      lambda = newImpulse - lam[2 * p];
      lam[2 * p] = newImpulse;
      TV2<T> P = lambda * c.normal;
      A.v = A.v - mA * P;
      A.w -= iA * cross(ra, P);
      B.v = B.v + mB * P;
      B.w += iB * cross(rb, P);
    }
    return;
  }
  TV2<T> aa = v2(lam[0], lam[2]);
  TV2<T> dv1 = ((B.v + cross(B.w, c.rb[0])) - A.v) - cross(A.w, c.ra[0]);
  TV2<T> dv2 = ((B.v + cross(B.w, c.rb[1])) - A.v) - cross(A.w, c.ra[1]);
  T vn1 = dot(dv1, c.normal);
  T vn2 = dot(dv2, c.normal);
  TV2<T> b = v2(vn1 - c.bias[0], vn2 - c.bias[1]);
  TV2<T> Ka = v2(c.K[0] * aa.x + c.K[2] * aa.y, c.K[1] * aa.x + c.K[3] * aa.y);
  b = b - Ka;
  // case 1: vn = 0
  TV2<T> x = -v2(c.NM[0] * b.x + c.NM[2] * b.y, c.NM[1] * b.x + c.NM[3] * b.y);
  if (x.x >= T(0.0) && x.y >= T(0.0)) {
    apply_normal_pair_t(A, B, k, c, x - aa);
    lam[0] = x.x;
    lam[2] = x.y;
    return;
  }
  // case 2: vn1 = 0 and x2 = 0
  x.x = -c.nm[0] * b.x;
  x.y = T(0.0);
  vn2 = c.K[1] * x.x + b.y;
  if (x.x >= T(0.0) && vn2 >= T(0.0)) {
    apply_normal_pair_t(A, B, k, c, x - aa);
    lam[0] = x.x;
    lam[2] = x.y;
    return;
  }
  // case 3: vn2 = 0 and x1 = 0
  x.x = T(0.0);
  x.y = -c.nm[1] * b.y;
  vn1 = c.K[2] * x.y + b.x;
  if (x.y >= T(0.0) && vn1 >= T(0.0)) {
    apply_normal_pair_t(A, B, k, c, x - aa);
    lam[0] = x.x;
    lam[2] = x.y;
    return;
  }
  // case 4: x1 = 0 and x2 = 0
  x.x = T(0.0);
  x.y = T(0.0);
  vn1 = b.x;
  vn2 = b.y;
  if (vn1 >= T(0.0) && vn2 >= T(0.0)) {
    apply_normal_pair_t(A, B, k, c, x - aa);
    lam[0] = x.x;
    lam[2] = x.y;
  }
  // no solution: the impulses stay as they are
}

// solve_position_constraints for one constraint, point by point; returns the smallest
// separation seen (src/private/dynamics/b2_contact_solver_private.rs:600-730). With the TOI
// Baumgarte factor and the non-TOI bodies' masses zeroed it is solve_toiposition_constraints
// (:733-822).
template <class T, class S>
B2G_HD T solve_position_t(const TContactConst<S>& k, const TGeom<T>& g, TPos<T>& A, TPos<T>& B,
                          double baumgarte = kBaumgarteD) {
  const T mA = T(k.mA), iA = T(k.iA), mB = T(k.mB), iB = T(k.iB);
  const TV2<T> lcA = v2(T(k.lcA[0]), T(k.lcA[1])), lcB = v2(T(k.lcB[0]), T(k.lcB[1]));
  T minSep = T(0.0);
  for (int p = 0; p < k.count; ++p) {
    TXf<T> xfA, xfB;
    xfA.q = rot(A.a);
    xfB.q = rot(B.a);
    xfA.p = A.c - mul(xfA.q, lcA);
    xfB.p = B.c - mul(xfB.q, lcB);
    TV2<T> normal, point;
    T separation;
    if (k.type == kCircles) {
      TV2<T> pA = mul(xfA, g.lpt);
      TV2<T> pB = mul(xfB, g.lp[0]);
      normal = pB - pA;
      normalize(&normal);
      point = T(0.5) * (pA + pB);
      separation = dot(pB - pA, normal) - T(k.rA) - T(k.rB);
    } else if (k.type == kFaceA) {
      normal = mul(xfA.q, g.ln);
      TV2<T> plane = mul(xfA, g.lpt);
      TV2<T> clip = mul(xfB, g.lp[p]);
      separation = dot(clip - plane, normal) - T(k.rA) - T(k.rB);
      point = clip;
    } else {
      normal = mul(xfB.q, g.ln);
      TV2<T> plane = mul(xfB, g.lpt);
      TV2<T> clip = mul(xfA, g.lp[p]);
      separation = dot(clip - plane, normal) - T(k.rA) - T(k.rB);
      point = clip;
      normal = -normal;
    }
    TV2<T> rA = point - A.c, rB = point - B.c;
    minSep = bmin(minSep, separation);
    T C = bclamp(T(baumgarte) * (separation + T(kLinearSlopD)), -T(kMaxLinearCorrectionD), T(0.0));
    T rnA = cross(rA, normal), rnB = cross(rB, normal);
    T K = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
    T impulse = K > T(0.0) ? -C / K : T(0.0);
    TV2<T> P = impulse * normal;
    A.c = A.c - mA * P;
    A.a -= iA * cross(rA, P);
    B.c = B.c + mB * P;
    B.a += iB * cross(rB, P);
  }
  return minSep;
}

}  // namespace b2g
