// b2g adjoint: vector-Jacobian products of the step's functions, shared by the GPU adjoint
// kernels (S = float) and host models (S = double).
//
// Each per-body or per-constraint function of the frozen step (the contact set, manifold types
// and feature ids, solve order and per-island position-iteration counts of a recorded forward
// run are fixed; everything continuous is differentiated) is written once, as a functor over a
// flat input vector x and output vector y, generic in the scalar. vjp() evaluates the functor
// on dual numbers L input lanes at a time and contracts the output tangents with the output
// cotangent: xbar += J(x)^T ybar. Branches (friction clamp, max(., 0), the block solver's case,
// the position clamp, translation/rotation limits, collide's feature choice) follow the primal
// values, so J is the Jacobian of the branch the forward run took.
#pragma once
#include "b2g_dual.h"
#include "b2g_step_t.h"

namespace b2g {

// one group of L input lanes, [base, base + L): xbar's entries of the group receive their
// contractions, the others are left alone (groups are independent: a kernel may run them on
// separate threads)
template <int NIN, int NOUT, int L, class S, class F>
B2G_HD void vjp_group(const S* x, const S* ybar, S* xbar, const F& f, int base) {
  Dual<L, S> xd[NIN], yd[NOUT];
  for (int i = 0; i < NIN; ++i) xd[i] = dual_var<L, S>(x[i], i - base);
  f(xd, yd);
  for (int j = 0; j < L && base + j < NIN; ++j) {
    S acc = S(0);
    for (int o = 0; o < NOUT; ++o) acc += ybar[o] * yd[o].d[j];
    xbar[base + j] += acc;
  }
}

template <int NIN, int NOUT, int L, class S, class F>
B2G_HD void vjp(const S* x, const S* ybar, S* xbar, const F& f) {
  for (int base = 0; base < NIN; base += L) vjp_group<NIN, NOUT, L>(x, ybar, xbar, f, base);
}

// ---- flat layouts ----------------------------------------------------------------------------
// position q = (c.x, c.y, a), velocity v = (v.x, v.y, w), geometry g = (ln, lpt, lp0, lp1),
// impulses lam = (normal0, tangent0, normal1, tangent1)
template <class T>
B2G_HD TPos<T> pos_of(const T* q) { TPos<T> p; p.c = v2(q[0], q[1]); p.a = q[2]; return p; }
template <class T>
B2G_HD TVel<T> vel_of(const T* v) { TVel<T> r; r.v = v2(v[0], v[1]); r.w = v[2]; return r; }
template <class T>
B2G_HD void put_pos(T* q, const TPos<T>& p) { q[0] = p.c.x; q[1] = p.c.y; q[2] = p.a; }
template <class T>
B2G_HD void put_vel(T* v, const TVel<T>& r) { v[0] = r.v.x; v[1] = r.v.y; v[2] = r.w; }
template <class T>
B2G_HD TGeom<T> geom_of(const T* g) {
  TGeom<T> r;
  r.ln = v2(g[0], g[1]);
  r.lpt = v2(g[2], g[3]);
  r.lp[0] = v2(g[4], g[5]);
  r.lp[1] = v2(g[6], g[7]);
  return r;
}

// collide with frozen features: manifold geometry of a contact from the two body positions
template <class S>
struct GeomFn {
  const ShapeV* sa;
  const ShapeV* sb;
  const S* lcA;
  const S* lcB;
  int* count;  // receives the primal point count (checked against the recorded one)
  uint32_t* ids;
  template <class T>
  B2G_HD void operator()(const T* x, T* y) const {
    TXf<T> xa = xf_of_t(pos_of(x), lcA), xb = xf_of_t(pos_of(x + 3), lcB);
    TManifold<T> m;
    evaluate(&m, *sa, xa, *sb, xb);
    y[0] = m.local_normal.x; y[1] = m.local_normal.y;
    y[2] = m.local_point.x; y[3] = m.local_point.y;
    y[4] = m.lp[0].x; y[5] = m.lp[0].y;
    y[6] = m.lp[1].x; y[7] = m.lp[1].y;
    if (count) *count = m.type | (m.count << 8);
    if (ids) { ids[0] = m.id[0]; ids[1] = m.id[1]; }
  }
};

// integrate velocities of a dynamic body: x = v (3) -> y = v* (3)
template <class S>
struct IntVelFn {
  TBodyMass<S> m;
  S fx, fy, torque, gx, gy, h;
  template <class T>
  B2G_HD void operator()(const T* x, T* y) const {
    TVel<T> v = vel_of(x);
    integrate_velocity_t(v, m, fx, fy, torque, gx, gy, h);
    put_vel(y, v);
  }
};

// integrate positions: x = (q, v) (6) -> y = (q', v') (6)
template <class S>
struct IntPosFn {
  S h;
  template <class T>
  B2G_HD void operator()(const T* x, T* y) const {
    TPos<T> p = pos_of(x);
    TVel<T> v = vel_of(x + 3);
    integrate_position_t(p, v, h);
    put_pos(y, p);
    put_vel(y + 3, v);
  }
};

// Input layout of the velocity-phase functors (31 values):
//   [0,3) vA  [3,6) vB  [6,10) lam  [10,13) qA  [13,16) qB  [16,24) g  [24] friction
//   [25,28) v*A  [28,31) v*B   (q and v* as at the start of the solve: they fix the constraint)
constexpr int kVelIn = 31;

template <class T, class S>
B2G_HD TVelC<T> velc_of(const T* x, const TContactConst<S>& k, bool block) {
  TGeom<T> g = geom_of(x + 16);
  return init_velocity_t(k, g, x[24], pos_of(x + 10), pos_of(x + 13), vel_of(x + 25), vel_of(x + 28), block);
}

// warm start: y = (vA', vB') (6)
template <class S>
struct WarmFn {
  TContactConst<S> k;
  bool block;
  template <class T>
  B2G_HD void operator()(const T* x, T* y) const {
    TVelC<T> c = velc_of(x, k, block);
    TVel<T> A = vel_of(x), B = vel_of(x + 3);
    T lam[4] = {x[6], x[7], x[8], x[9]};
    warm_start_t(c, k, lam, A, B);
    put_vel(y, A);
    put_vel(y + 3, B);
  }
};

// one velocity iteration of one constraint: y = (vA', vB', lam') (10)
template <class S>
struct VelFn {
  TContactConst<S> k;
  bool block;
  template <class T>
  B2G_HD void operator()(const T* x, T* y) const {
    TVelC<T> c = velc_of(x, k, block);
    TVel<T> A = vel_of(x), B = vel_of(x + 3);
    T lam[4] = {x[6], x[7], x[8], x[9]};
    solve_velocity_t(c, k, lam, A, B, block);
    put_vel(y, A);
    put_vel(y + 3, B);
    for (int i = 0; i < 4; ++i) y[6 + i] = lam[i];
  }
};

// one position iteration of one constraint: x = (qA, qB, g) (14) -> y = (qA', qB') (6)
template <class S>
struct PosFn {
  TContactConst<S> k;
  template <class T>
  B2G_HD void operator()(const T* x, T* y) const {
    TPos<T> A = pos_of(x), B = pos_of(x + 3);
    TGeom<T> g = geom_of(x + 6);
    solve_position_t(k, g, A, B);
    put_pos(y, A);
    put_pos(y + 3, B);
  }
};

}  // namespace b2g
