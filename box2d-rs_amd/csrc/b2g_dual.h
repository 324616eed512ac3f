// b2g: forward-mode dual numbers for the adjoint.
//
// Dual<N, S> carries a primal value of scalar type S and N tangent lanes. Evaluating one of the
// templated step functions (collide, constraint solves, integration) on Duals seeded with unit
// tangents gives N columns of its Jacobian at the primal point; the adjoint contracts those
// with the output cotangents (a vector-Jacobian product assembled from local Jacobians).
// Comparisons look at the primal value only, so every branch (clamps, the block solver's case,
// feature selection in collide) is the one the primal evaluation takes: the derivative is that
// of the frozen active set.
#pragma once
#include "b2g_math.h"

namespace b2g {

template <int N, class S>
struct Dual {
  S v;
  S d[N];
  B2G_HD Dual() : v(S(0)) {
    for (int i = 0; i < N; ++i) d[i] = S(0);
  }
  B2G_HD Dual(double x) : v(S(x)) {  // constants (the generic code writes T(literal))
    for (int i = 0; i < N; ++i) d[i] = S(0);
  }
  B2G_HD Dual(float x) : v(S(x)) {
    for (int i = 0; i < N; ++i) d[i] = S(0);
  }
  B2G_HD Dual& operator+=(const Dual& b) { *this = *this + b; return *this; }
  B2G_HD Dual& operator-=(const Dual& b) { *this = *this - b; return *this; }
  B2G_HD Dual& operator*=(const Dual& b) { *this = *this * b; return *this; }
};

template <int N, class S>
B2G_HD Dual<N, S> dual_var(S v, int lane) {
  Dual<N, S> r;
  r.v = v;
  if (lane >= 0 && lane < N) r.d[lane] = S(1);
  return r;
}

template <int N, class S>
B2G_HD Dual<N, S> operator+(const Dual<N, S>& a, const Dual<N, S>& b) {
  Dual<N, S> r;
  r.v = a.v + b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N, class S>
B2G_HD Dual<N, S> operator-(const Dual<N, S>& a, const Dual<N, S>& b) {
  Dual<N, S> r;
  r.v = a.v - b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N, class S>
B2G_HD Dual<N, S> operator-(const Dual<N, S>& a) {
  Dual<N, S> r;
  r.v = -a.v;
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N, class S>
B2G_HD Dual<N, S> operator*(const Dual<N, S>& a, const Dual<N, S>& b) {
  Dual<N, S> r;
  r.v = a.v * b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N, class S>
B2G_HD Dual<N, S> operator/(const Dual<N, S>& a, const Dual<N, S>& b) {
  Dual<N, S> r;
  r.v = a.v / b.v;
  const S inv = S(1) / b.v;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
  return r;
}
template <int N, class S>
B2G_HD bool operator<(const Dual<N, S>& a, const Dual<N, S>& b) { return a.v < b.v; }
template <int N, class S>
B2G_HD bool operator>(const Dual<N, S>& a, const Dual<N, S>& b) { return a.v > b.v; }
template <int N, class S>
B2G_HD bool operator<=(const Dual<N, S>& a, const Dual<N, S>& b) { return a.v <= b.v; }
template <int N, class S>
B2G_HD bool operator>=(const Dual<N, S>& a, const Dual<N, S>& b) { return a.v >= b.v; }
template <int N, class S>
B2G_HD bool operator==(const Dual<N, S>& a, const Dual<N, S>& b) { return a.v == b.v; }
template <int N, class S>
B2G_HD bool operator!=(const Dual<N, S>& a, const Dual<N, S>& b) { return a.v != b.v; }

template <int N, class S>
B2G_HD S primal(const Dual<N, S>& a) { return a.v; }

template <int N, class S>
B2G_HD Dual<N, S> tsqrt(const Dual<N, S>& a) {
  Dual<N, S> r;
  r.v = tsqrt(a.v);
  const S k = r.v > S(0) ? S(0.5) / r.v : S(0);
  for (int i = 0; i < N; ++i) r.d[i] = k * a.d[i];
  return r;
}

// rotation of a Dual angle: the primal is the scalar type's own rot(), the tangent follows
// d(sin a) = cos a da, d(cos a) = -sin a da
template <int N, class S>
B2G_HD TRot<Dual<N, S>> rot(const Dual<N, S>& a) {
  TRot<S> q = rot(a.v);
  TRot<Dual<N, S>> r;
  r.s.v = q.s;
  r.c.v = q.c;
  for (int i = 0; i < N; ++i) {
    r.s.d[i] = q.c * a.d[i];
    r.c.d[i] = -q.s * a.d[i];
  }
  return r;
}

}  // namespace b2g
