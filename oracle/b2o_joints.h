// Oracle (test infrastructure, never part of the product): the distance, revolute, weld and
// prismatic joint solvers of the reference, restated in operation order.
//   distance joint  src/private/dynamics/joints/b2_distance_joint.rs:23-330
//   revolute joint  src/private/dynamics/joints/b2_revolute_joint.rs:60-330, construction
//                   src/joints/b2_revolute_joint.rs:253-287
//   weld joint      src/private/dynamics/joints/b2_weld_joint.rs:23-300
//   prismatic joint src/private/dynamics/joints/b2_prismatic_joint.rs:40-460
//   B2Mat22::solve  src/b2_math.rs:278-293; B2Mat33 solve33 / solve22 / get_inverse22 /
//   get_sym_inverse33 src/private/common/b2_math.rs:5-71, b2_mul_mat33 / b2_mul22 src/b2_math.rs:601-608
#pragma once

#include "b2o_world.h"

namespace b2o {

static const real kAngularSlop = real(2.0) / real(180.0) * kPi;            // src/b2_common.rs:43
static const real kMaxAngularCorrection = real(8.0) / real(180.0) * kPi;   // src/b2_common.rs:64

inline Vec2 mat22_solve(const Mat22& k, Vec2 b) {
  const real a11 = k.ex.x, a12 = k.ey.x, a21 = k.ex.y, a22 = k.ey.y;
  real det = a11 * a22 - a12 * a21;
  if (det != 0) det = real(1.0) / det;
  return Vec2(det * (a22 * b.x - a12 * b.y), det * (a11 * b.y - a21 * b.x));
}

// the joint's bodies as the island solver holds them
struct JointSolverData {
  TimeStep step;
  Position* positions;
  Velocity* velocities;
};

inline void joint_bind(Joint* j) {
  j->index_a = j->a->island_index;
  j->index_b = j->b->island_index;
  j->local_center_a = j->a->sweep.local_center;
  j->local_center_b = j->b->sweep.local_center;
  j->inv_mass_a = j->a->inv_mass;
  j->inv_mass_b = j->b->inv_mass;
  j->inv_ia = j->a->inv_I;
  j->inv_ib = j->b->inv_I;
}

// ---- distance joint -----------------------------------------------------------------------
inline void distance_init_velocity(Joint* j, const JointSolverData& d) {
  joint_bind(j);
  Vec2 cA = d.positions[j->index_a].c;
  real aA = d.positions[j->index_a].a;
  Vec2 vA = d.velocities[j->index_a].v;
  real wA = d.velocities[j->index_a].w;
  Vec2 cB = d.positions[j->index_b].c;
  real aB = d.positions[j->index_b].a;
  Vec2 vB = d.velocities[j->index_b].v;
  real wB = d.velocities[j->index_b].w;
  Rot qA, qB;
  qA.set(aA);
  qB.set(aB);
  j->rA = mul(qA, j->local_anchor_a - j->local_center_a);
  j->rB = mul(qB, j->local_anchor_b - j->local_center_b);
  j->u = cB + j->rB - cA - j->rA;
  j->current_length = j->u.length();
  if (j->current_length > kLinearSlop) {
    j->u *= real(1.0) / j->current_length;
  } else {
    j->u = Vec2(0, 0);
    j->mass = 0;
    j->impulse = 0;
    j->lower_impulse = 0;
    j->upper_impulse = 0;
  }
  const real crAu = cross(j->rA, j->u);
  const real crBu = cross(j->rB, j->u);
  real inv_mass = j->inv_mass_a + j->inv_ia * crAu * crAu + j->inv_mass_b + j->inv_ib * crBu * crBu;
  j->mass = inv_mass != 0 ? real(1.0) / inv_mass : real(0);
  if (j->stiffness > 0 && j->min_length < j->max_length) {
    const real c = j->current_length - j->length;
    const real dmp = j->damping;
    const real k = j->stiffness;
    const real h = d.step.dt;
    j->gamma = h * (dmp + h * k);
    j->gamma = j->gamma != 0 ? real(1.0) / j->gamma : real(0);
    j->bias = c * h * k * j->gamma;
    inv_mass += j->gamma;
    j->soft_mass = inv_mass != 0 ? real(1.0) / inv_mass : real(0);
  } else {
    j->gamma = 0;
    j->bias = 0;
    j->soft_mass = j->mass;
  }
  if (d.step.warm_starting) {
    j->impulse *= d.step.dt_ratio;
    j->lower_impulse *= d.step.dt_ratio;
    j->upper_impulse *= d.step.dt_ratio;
    const Vec2 P = (j->impulse + j->lower_impulse - j->upper_impulse) * j->u;
    vA -= j->inv_mass_a * P;
    wA -= j->inv_ia * cross(j->rA, P);
    vB += j->inv_mass_b * P;
    wB += j->inv_ib * cross(j->rB, P);
  } else {
    j->impulse = 0;
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline void distance_apply(Joint* j, real impulse, Vec2& vA, real& wA, Vec2& vB, real& wB) {
  const Vec2 P = impulse * j->u;
  vA -= j->inv_mass_a * P;
  wA -= j->inv_ia * cross(j->rA, P);
  vB += j->inv_mass_b * P;
  wB += j->inv_ib * cross(j->rB, P);
}

inline void distance_solve_velocity(Joint* j, const JointSolverData& d) {
  Vec2 vA = d.velocities[j->index_a].v;
  real wA = d.velocities[j->index_a].w;
  Vec2 vB = d.velocities[j->index_b].v;
  real wB = d.velocities[j->index_b].w;
  if (j->min_length < j->max_length) {
    if (j->stiffness > 0) {
      const Vec2 vpA = vA + cross(wA, j->rA);
      const Vec2 vpB = vB + cross(wB, j->rB);
      const real Cdot = dot(j->u, vpB - vpA);
      const real impulse = -j->soft_mass * (Cdot + j->bias + j->gamma * j->impulse);
      j->impulse += impulse;
      distance_apply(j, impulse, vA, wA, vB, wB);
    }
    {  // lower
      const real C = j->current_length - j->min_length;
      const real bias = bmax(real(0), C) * d.step.inv_dt;
      const Vec2 vpA = vA + cross(wA, j->rA);
      const Vec2 vpB = vB + cross(wB, j->rB);
      const real Cdot = dot(j->u, vpB - vpA);
      real impulse = -j->mass * (Cdot + bias);
      const real old = j->lower_impulse;
      j->lower_impulse = bmax(real(0), j->lower_impulse + impulse);
      impulse = j->lower_impulse - old;
      distance_apply(j, impulse, vA, wA, vB, wB);
    }
    {  // upper
      const real C = j->max_length - j->current_length;
      const real bias = bmax(real(0), C) * d.step.inv_dt;
      const Vec2 vpA = vA + cross(wA, j->rA);
      const Vec2 vpB = vB + cross(wB, j->rB);
      const real Cdot = dot(j->u, vpA - vpB);
      real impulse = -j->mass * (Cdot + bias);
      const real old = j->upper_impulse;
      j->upper_impulse = bmax(real(0), j->upper_impulse + impulse);
      impulse = j->upper_impulse - old;
      distance_apply(j, -impulse, vA, wA, vB, wB);
    }
  } else {
    const Vec2 vpA = vA + cross(wA, j->rA);
    const Vec2 vpB = vB + cross(wB, j->rB);
    const real Cdot = dot(j->u, vpB - vpA);
    const real impulse = -j->mass * Cdot;
    j->impulse += impulse;
    distance_apply(j, impulse, vA, wA, vB, wB);
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline bool distance_solve_position(Joint* j, const JointSolverData& d) {
  Vec2 cA = d.positions[j->index_a].c;
  real aA = d.positions[j->index_a].a;
  Vec2 cB = d.positions[j->index_b].c;
  real aB = d.positions[j->index_b].a;
  Rot qA, qB;
  qA.set(aA);
  qB.set(aB);
  const Vec2 rA = mul(qA, j->local_anchor_a - j->local_center_a);
  const Vec2 rB = mul(qB, j->local_anchor_b - j->local_center_b);
  Vec2 u = cB + rB - cA - rA;
  const real length = u.normalize();
  real C;
  if (j->min_length == j->max_length) C = length - j->min_length;
  else if (length < j->min_length) C = length - j->min_length;
  else if (j->max_length < length) C = length - j->max_length;
  else return true;
  const real impulse = -j->mass * C;
  const Vec2 P = impulse * u;
  cA -= j->inv_mass_a * P;
  aA -= j->inv_ia * cross(rA, P);
  cB += j->inv_mass_b * P;
  aB += j->inv_ib * cross(rB, P);
  d.positions[j->index_a].c = cA;
  d.positions[j->index_a].a = aA;
  d.positions[j->index_b].c = cB;
  d.positions[j->index_b].a = aB;
  return babs(C) < kLinearSlop;
}

// ---- revolute joint -----------------------------------------------------------------------
inline void revolute_init_velocity(Joint* j, const JointSolverData& d) {
  joint_bind(j);
  const real aA = d.positions[j->index_a].a;
  const real aB = d.positions[j->index_b].a;
  Vec2 vA = d.velocities[j->index_a].v;
  real wA = d.velocities[j->index_a].w;
  Vec2 vB = d.velocities[j->index_b].v;
  real wB = d.velocities[j->index_b].w;
  Rot qA, qB;
  qA.set(aA);
  qB.set(aB);
  j->rA = mul(qA, j->local_anchor_a - j->local_center_a);
  j->rB = mul(qB, j->local_anchor_b - j->local_center_b);
  const real mA = j->inv_mass_a, mB = j->inv_mass_b, iA = j->inv_ia, iB = j->inv_ib;
  j->K.ex.x = mA + mB + j->rA.y * j->rA.y * iA + j->rB.y * j->rB.y * iB;
  j->K.ey.x = -j->rA.y * j->rA.x * iA - j->rB.y * j->rB.x * iB;
  j->K.ex.y = j->K.ey.x;
  j->K.ey.y = mA + mB + j->rA.x * j->rA.x * iA + j->rB.x * j->rB.x * iB;
  j->axial_mass = iA + iB;
  bool fixed_rotation;
  if (j->axial_mass > 0) {
    j->axial_mass = real(1.0) / j->axial_mass;
    fixed_rotation = false;
  } else {
    fixed_rotation = true;
  }
  j->angle = aB - aA - j->reference_angle;
  if (!j->enable_limit || fixed_rotation) {
    j->lower_impulse = 0;
    j->upper_impulse = 0;
  }
  if (!j->enable_motor || fixed_rotation) j->motor_impulse = 0;
  if (d.step.warm_starting) {
    j->impulse2 *= d.step.dt_ratio;
    j->motor_impulse *= d.step.dt_ratio;
    j->lower_impulse *= d.step.dt_ratio;
    j->upper_impulse *= d.step.dt_ratio;
    const real axial_impulse = j->motor_impulse + j->lower_impulse - j->upper_impulse;
    const Vec2 P(j->impulse2.x, j->impulse2.y);
    vA -= mA * P;
    wA -= iA * (cross(j->rA, P) + axial_impulse);
    vB += mB * P;
    wB += iB * (cross(j->rB, P) + axial_impulse);
  } else {
    j->impulse2 = Vec2(0, 0);
    j->motor_impulse = 0;
    j->lower_impulse = 0;
    j->upper_impulse = 0;
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline void revolute_solve_velocity(Joint* j, const JointSolverData& d) {
  Vec2 vA = d.velocities[j->index_a].v;
  real wA = d.velocities[j->index_a].w;
  Vec2 vB = d.velocities[j->index_b].v;
  real wB = d.velocities[j->index_b].w;
  const real mA = j->inv_mass_a, mB = j->inv_mass_b, iA = j->inv_ia, iB = j->inv_ib;
  const bool fixed_rotation = iA + iB == 0;
  if (j->enable_motor && !fixed_rotation) {
    const real Cdot = wB - wA - j->motor_speed;
    real impulse = -j->axial_mass * Cdot;
    const real old = j->motor_impulse;
    const real max_impulse = d.step.dt * j->max_motor_torque;
    j->motor_impulse = bclamp(j->motor_impulse + impulse, -max_impulse, max_impulse);
    impulse = j->motor_impulse - old;
    wA -= iA * impulse;
    wB += iB * impulse;
  }
  if (j->enable_limit && !fixed_rotation) {
    {
      const real C = j->angle - j->lower_angle;
      const real Cdot = wB - wA;
      real impulse = -j->axial_mass * (Cdot + bmax(C, real(0)) * d.step.inv_dt);
      const real old = j->lower_impulse;
      j->lower_impulse = bmax(j->lower_impulse + impulse, real(0));
      impulse = j->lower_impulse - old;
      wA -= iA * impulse;
      wB += iB * impulse;
    }
    {
      const real C = j->upper_angle - j->angle;
      const real Cdot = wA - wB;
      real impulse = -j->axial_mass * (Cdot + bmax(C, real(0)) * d.step.inv_dt);
      const real old = j->upper_impulse;
      j->upper_impulse = bmax(j->upper_impulse + impulse, real(0));
      impulse = j->upper_impulse - old;
      wA += iA * impulse;
      wB -= iB * impulse;
    }
  }
  {
    const Vec2 Cdot = vB + cross(wB, j->rB) - vA - cross(wA, j->rA);
    const Vec2 impulse = mat22_solve(j->K, -Cdot);
    j->impulse2.x += impulse.x;
    j->impulse2.y += impulse.y;
    vA -= mA * impulse;
    wA -= iA * cross(j->rA, impulse);
    vB += mB * impulse;
    wB += iB * cross(j->rB, impulse);
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline bool revolute_solve_position(Joint* j, const JointSolverData& d) {
  Vec2 cA = d.positions[j->index_a].c;
  real aA = d.positions[j->index_a].a;
  Vec2 cB = d.positions[j->index_b].c;
  real aB = d.positions[j->index_b].a;
  Rot qA, qB;
  qA.set(aA);
  qB.set(aB);
  real angular_error = 0;
  real position_error;
  const bool fixed_rotation = j->inv_ia + j->inv_ib == 0;
  if (j->enable_limit && !fixed_rotation) {
    const real angle = aB - aA - j->reference_angle;
    real C = 0;
    if (babs(j->upper_angle - j->lower_angle) < real(2.0) * kAngularSlop) {
      C = bclamp(angle - j->lower_angle, -kMaxAngularCorrection, kMaxAngularCorrection);
    } else if (angle <= j->lower_angle) {
      C = bclamp(angle - j->lower_angle + kAngularSlop, -kMaxAngularCorrection, real(0));
    } else if (angle >= j->upper_angle) {
      C = bclamp(angle - j->upper_angle - kAngularSlop, real(0), kMaxAngularCorrection);
    }
    const real limit_impulse = -j->axial_mass * C;
    aA -= j->inv_ia * limit_impulse;
    aB += j->inv_ib * limit_impulse;
    angular_error = babs(C);
  }
  {
    qA.set(aA);
    qB.set(aB);
    const Vec2 rA = mul(qA, j->local_anchor_a - j->local_center_a);
    const Vec2 rB = mul(qB, j->local_anchor_b - j->local_center_b);
    const Vec2 C = cB + rB - cA - rA;
    position_error = C.length();
    const real mA = j->inv_mass_a, mB = j->inv_mass_b, iA = j->inv_ia, iB = j->inv_ib;
    Mat22 K;
    K.ex.x = mA + mB + iA * rA.y * rA.y + iB * rB.y * rB.y;
    K.ex.y = -iA * rA.x * rA.y - iB * rB.x * rB.y;
    K.ey.x = K.ex.y;
    K.ey.y = mA + mB + iA * rA.x * rA.x + iB * rB.x * rB.x;
    const Vec2 impulse = -mat22_solve(K, C);
    cA -= mA * impulse;
    aA -= iA * cross(rA, impulse);
    cB += mB * impulse;
    aB += iB * cross(rB, impulse);
  }
  d.positions[j->index_a].c = cA;
  d.positions[j->index_a].a = aA;
  d.positions[j->index_b].c = cB;
  d.positions[j->index_b].a = aB;
  return position_error <= kLinearSlop && angular_error <= kAngularSlop;
}

// ---- 3x3 helpers: columns ex, ey, ez --------------------------------------------------------
struct Vec3 {
  real x = 0, y = 0, z = 0;
  Vec3() = default;
  Vec3(real x_, real y_, real z_) : x(x_), y(y_), z(z_) {}
};
struct Mat33 {
  Vec3 ex, ey, ez;
};
inline real dot3(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross3(Vec3 a, Vec3 b) {
  return Vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline Vec3 solve33(const Mat33& k, Vec3 b) {
  real det = dot3(k.ex, cross3(k.ey, k.ez));
  if (det != 0) det = real(1.0) / det;
  return Vec3(det * dot3(b, cross3(k.ey, k.ez)), det * dot3(k.ex, cross3(b, k.ez)), det * dot3(k.ex, cross3(k.ey, b)));
}
inline Vec2 solve22(const Mat33& k, Vec2 b) {
  const real a11 = k.ex.x, a12 = k.ey.x, a21 = k.ex.y, a22 = k.ey.y;
  real det = a11 * a22 - a12 * a21;
  if (det != 0) det = real(1.0) / det;
  return Vec2(det * (a22 * b.x - a12 * b.y), det * (a11 * b.y - a21 * b.x));
}
inline void inverse22(const Mat33& k, Mat33* m) {
  const real a = k.ex.x, b = k.ey.x, c = k.ex.y, d = k.ey.y;
  real det = a * d - b * c;
  if (det != 0) det = real(1.0) / det;
  m->ex.x = det * d;
  m->ey.x = -det * b;
  m->ex.z = 0;
  m->ex.y = -det * c;
  m->ey.y = det * a;
  m->ey.z = 0;
  m->ez.x = 0;
  m->ez.y = 0;
  m->ez.z = 0;
}
inline void sym_inverse33(const Mat33& k, Mat33* m) {
  real det = dot3(k.ex, cross3(k.ey, k.ez));
  if (det != 0) det = real(1.0) / det;
  const real a11 = k.ex.x, a12 = k.ey.x, a13 = k.ez.x, a22 = k.ey.y, a23 = k.ez.y, a33 = k.ez.z;
  m->ex.x = det * (a22 * a33 - a23 * a23);
  m->ex.y = det * (a13 * a23 - a12 * a33);
  m->ex.z = det * (a12 * a23 - a13 * a22);
  m->ey.x = m->ex.y;
  m->ey.y = det * (a11 * a33 - a13 * a13);
  m->ey.z = det * (a13 * a12 - a11 * a23);
  m->ez.x = m->ex.z;
  m->ez.y = m->ey.z;
  m->ez.z = det * (a11 * a22 - a12 * a12);
}
// v.x * ex + v.y * ey + v.z * ez
inline Vec3 mul33(const Mat33& a, Vec3 v) {
  return Vec3(v.x * a.ex.x + v.y * a.ey.x + v.z * a.ez.x, v.x * a.ex.y + v.y * a.ey.y + v.z * a.ez.y,
              v.x * a.ex.z + v.y * a.ey.z + v.z * a.ez.z);
}
inline Vec2 mul22(const Mat33& a, Vec2 v) { return Vec2(a.ex.x * v.x + a.ey.x * v.y, a.ex.y * v.x + a.ey.y * v.y); }
inline Mat33 weld_k(real mA, real mB, real iA, real iB, Vec2 rA, Vec2 rB) {
  Mat33 k;
  k.ex.x = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
  k.ey.x = -rA.y * rA.x * iA - rB.y * rB.x * iB;
  k.ez.x = -rA.y * iA - rB.y * iB;
  k.ex.y = k.ey.x;
  k.ey.y = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
  k.ez.y = rA.x * iA + rB.x * iB;
  k.ex.z = k.ez.x;
  k.ey.z = k.ez.y;
  k.ez.z = iA + iB;
  return k;
}
inline Mat33 mass_of(const Joint* j) {
  Mat33 m;
  m.ex = Vec3(j->mass33[0][0], j->mass33[0][1], j->mass33[0][2]);
  m.ey = Vec3(j->mass33[1][0], j->mass33[1][1], j->mass33[1][2]);
  m.ez = Vec3(j->mass33[2][0], j->mass33[2][1], j->mass33[2][2]);
  return m;
}
inline void set_mass(Joint* j, const Mat33& m) {
  const Vec3* cols[3] = {&m.ex, &m.ey, &m.ez};
  for (int c = 0; c < 3; ++c) {
    j->mass33[c][0] = cols[c]->x;
    j->mass33[c][1] = cols[c]->y;
    j->mass33[c][2] = cols[c]->z;
  }
}

// ---- weld joint ----------------------------------------------------------------------------
inline void weld_init_velocity(Joint* j, const JointSolverData& d) {
  joint_bind(j);
  const real aA = d.positions[j->index_a].a;
  const real aB = d.positions[j->index_b].a;
  Vec2 vA = d.velocities[j->index_a].v;
  real wA = d.velocities[j->index_a].w;
  Vec2 vB = d.velocities[j->index_b].v;
  real wB = d.velocities[j->index_b].w;
  Rot qA, qB;
  qA.set(aA);
  qB.set(aB);
  j->rA = mul(qA, j->local_anchor_a - j->local_center_a);
  j->rB = mul(qB, j->local_anchor_b - j->local_center_b);
  const real mA = j->inv_mass_a, mB = j->inv_mass_b, iA = j->inv_ia, iB = j->inv_ib;
  const Mat33 K = weld_k(mA, mB, iA, iB, j->rA, j->rB);
  Mat33 M;
  if (j->stiffness > 0) {
    inverse22(K, &M);
    real inv_m = iA + iB;
    const real C = aB - aA - j->reference_angle;
    const real dmp = j->damping, k = j->stiffness, h = d.step.dt;
    j->gamma = h * (dmp + h * k);
    j->gamma = j->gamma != 0 ? real(1.0) / j->gamma : real(0);
    j->bias = C * h * k * j->gamma;
    inv_m += j->gamma;
    M.ez.z = inv_m != 0 ? real(1.0) / inv_m : real(0);
  } else if (K.ez.z == 0) {
    inverse22(K, &M);
    j->gamma = 0;
    j->bias = 0;
  } else {
    sym_inverse33(K, &M);
    j->gamma = 0;
    j->bias = 0;
  }
  set_mass(j, M);
  if (d.step.warm_starting) {
    j->impulse3[0] *= d.step.dt_ratio;
    j->impulse3[1] *= d.step.dt_ratio;
    j->impulse3[2] *= d.step.dt_ratio;
    const Vec2 P(j->impulse3[0], j->impulse3[1]);
    vA -= mA * P;
    wA -= iA * (cross(j->rA, P) + j->impulse3[2]);
    vB += mB * P;
    wB += iB * (cross(j->rB, P) + j->impulse3[2]);
  } else {
    j->impulse3[0] = j->impulse3[1] = j->impulse3[2] = 0;
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline void weld_solve_velocity(Joint* j, const JointSolverData& d) {
  Vec2 vA = d.velocities[j->index_a].v;
  real wA = d.velocities[j->index_a].w;
  Vec2 vB = d.velocities[j->index_b].v;
  real wB = d.velocities[j->index_b].w;
  const real mA = j->inv_mass_a, mB = j->inv_mass_b, iA = j->inv_ia, iB = j->inv_ib;
  const Mat33 M = mass_of(j);
  if (j->stiffness > 0) {
    const real Cdot2 = wB - wA;
    const real impulse2 = -M.ez.z * (Cdot2 + j->bias + j->gamma * j->impulse3[2]);
    j->impulse3[2] += impulse2;
    wA -= iA * impulse2;
    wB += iB * impulse2;
    const Vec2 Cdot1 = vB + cross(wB, j->rB) - vA - cross(wA, j->rA);
    const Vec2 impulse1 = -mul22(M, Cdot1);
    j->impulse3[0] += impulse1.x;
    j->impulse3[1] += impulse1.y;
    const Vec2 P = impulse1;
    vA -= mA * P;
    wA -= iA * cross(j->rA, P);
    vB += mB * P;
    wB += iB * cross(j->rB, P);
  } else {
    const Vec2 Cdot1 = vB + cross(wB, j->rB) - vA - cross(wA, j->rA);
    const real Cdot2 = wB - wA;
    const Vec3 Cdot(Cdot1.x, Cdot1.y, Cdot2);
    const Vec3 m = mul33(M, Cdot);
    const Vec3 impulse(-m.x, -m.y, -m.z);
    j->impulse3[0] += impulse.x;
    j->impulse3[1] += impulse.y;
    j->impulse3[2] += impulse.z;
    const Vec2 P(impulse.x, impulse.y);
    vA -= mA * P;
    wA -= iA * (cross(j->rA, P) + impulse.z);
    vB += mB * P;
    wB += iB * (cross(j->rB, P) + impulse.z);
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline bool weld_solve_position(Joint* j, const JointSolverData& d) {
  Vec2 cA = d.positions[j->index_a].c;
  real aA = d.positions[j->index_a].a;
  Vec2 cB = d.positions[j->index_b].c;
  real aB = d.positions[j->index_b].a;
  Rot qA, qB;
  qA.set(aA);
  qB.set(aB);
  const real mA = j->inv_mass_a, mB = j->inv_mass_b, iA = j->inv_ia, iB = j->inv_ib;
  const Vec2 rA = mul(qA, j->local_anchor_a - j->local_center_a);
  const Vec2 rB = mul(qB, j->local_anchor_b - j->local_center_b);
  real position_error, angular_error;
  const Mat33 K = weld_k(mA, mB, iA, iB, rA, rB);
  if (j->stiffness > 0) {
    const Vec2 C1 = cB + rB - cA - rA;
    position_error = C1.length();
    angular_error = 0;
    const Vec2 P = -solve22(K, C1);
    cA -= mA * P;
    aA -= iA * cross(rA, P);
    cB += mB * P;
    aB += iB * cross(rB, P);
  } else {
    const Vec2 C1 = cB + rB - cA - rA;
    const real C2 = aB - aA - j->reference_angle;
    position_error = C1.length();
    angular_error = babs(C2);
    const Vec3 C(C1.x, C1.y, C2);
    Vec3 impulse;
    if (K.ez.z > 0) {
      const Vec3 s = solve33(K, C);
      impulse = Vec3(-s.x, -s.y, -s.z);
    } else {
      const Vec2 i2 = -solve22(K, C1);
      impulse = Vec3(i2.x, i2.y, 0);
    }
    const Vec2 P(impulse.x, impulse.y);
    cA -= mA * P;
    aA -= iA * (cross(rA, P) + impulse.z);
    cB += mB * P;
    aB += iB * (cross(rB, P) + impulse.z);
  }
  d.positions[j->index_a].c = cA;
  d.positions[j->index_a].a = aA;
  d.positions[j->index_b].c = cB;
  d.positions[j->index_b].a = aB;
  return position_error <= kLinearSlop && angular_error <= kAngularSlop;
}

// ---- prismatic joint ------------------------------------------------------------------------
inline void prismatic_push(Joint* j, Vec2 P, real LA, real LB, Vec2& vA, real& wA, Vec2& vB, real& wB) {
  vA -= j->inv_mass_a * P;
  wA -= j->inv_ia * LA;
  vB += j->inv_mass_b * P;
  wB += j->inv_ib * LB;
}

inline void prismatic_init_velocity(Joint* j, const JointSolverData& d) {
  joint_bind(j);
  const Vec2 cA = d.positions[j->index_a].c;
  const real aA = d.positions[j->index_a].a;
  Vec2 vA = d.velocities[j->index_a].v;
  real wA = d.velocities[j->index_a].w;
  const Vec2 cB = d.positions[j->index_b].c;
  const real aB = d.positions[j->index_b].a;
  Vec2 vB = d.velocities[j->index_b].v;
  real wB = d.velocities[j->index_b].w;
  Rot qA, qB;
  qA.set(aA);
  qB.set(aB);
  const Vec2 rA = mul(qA, j->local_anchor_a - j->local_center_a);
  const Vec2 rB = mul(qB, j->local_anchor_b - j->local_center_b);
  const Vec2 dd = (cB - cA) + rB - rA;
  const real mA = j->inv_mass_a, mB = j->inv_mass_b, iA = j->inv_ia, iB = j->inv_ib;
  {
    j->axis = mul(qA, j->local_xaxis_a);
    j->a1 = cross(dd + rA, j->axis);
    j->a2 = cross(rB, j->axis);
    j->axial_mass = mA + mB + iA * j->a1 * j->a1 + iB * j->a2 * j->a2;
    if (j->axial_mass > 0) j->axial_mass = real(1.0) / j->axial_mass;
  }
  {
    j->perp = mul(qA, j->local_yaxis_a);
    j->s1 = cross(dd + rA, j->perp);
    j->s2 = cross(rB, j->perp);
    const real k11 = mA + mB + iA * j->s1 * j->s1 + iB * j->s2 * j->s2;
    const real k12 = iA * j->s1 + iB * j->s2;
    real k22 = iA + iB;
    if (k22 == 0) k22 = 1;
    j->K.ex = Vec2(k11, k12);
    j->K.ey = Vec2(k12, k22);
  }
  if (j->enable_limit) {
    j->translation = dot(j->axis, dd);
  } else {
    j->lower_impulse = 0;
    j->upper_impulse = 0;
  }
  if (!j->enable_motor) j->motor_impulse = 0;
  if (d.step.warm_starting) {
    j->impulse2 *= d.step.dt_ratio;
    j->motor_impulse *= d.step.dt_ratio;
    j->lower_impulse *= d.step.dt_ratio;
    j->upper_impulse *= d.step.dt_ratio;
    const real axial_impulse = j->motor_impulse + j->lower_impulse - j->upper_impulse;
    const Vec2 P = j->impulse2.x * j->perp + axial_impulse * j->axis;
    const real LA = j->impulse2.x * j->s1 + j->impulse2.y + axial_impulse * j->a1;
    const real LB = j->impulse2.x * j->s2 + j->impulse2.y + axial_impulse * j->a2;
    prismatic_push(j, P, LA, LB, vA, wA, vB, wB);
  } else {
    j->impulse2 = Vec2(0, 0);
    j->motor_impulse = 0;
    j->lower_impulse = 0;
    j->upper_impulse = 0;
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline void prismatic_solve_velocity(Joint* j, const JointSolverData& d) {
  Vec2 vA = d.velocities[j->index_a].v;
  real wA = d.velocities[j->index_a].w;
  Vec2 vB = d.velocities[j->index_b].v;
  real wB = d.velocities[j->index_b].w;
  if (j->enable_motor) {
    const real Cdot = dot(j->axis, vB - vA) + j->a2 * wB - j->a1 * wA;
    real impulse = j->axial_mass * (j->motor_speed - Cdot);
    const real old = j->motor_impulse;
    const real max_impulse = d.step.dt * j->max_motor_force;
    j->motor_impulse = bclamp(j->motor_impulse + impulse, -max_impulse, max_impulse);
    impulse = j->motor_impulse - old;
    prismatic_push(j, impulse * j->axis, impulse * j->a1, impulse * j->a2, vA, wA, vB, wB);
  }
  if (j->enable_limit) {
    {
      const real C = j->translation - j->lower_translation;
      const real Cdot = dot(j->axis, vB - vA) + j->a2 * wB - j->a1 * wA;
      real impulse = -j->axial_mass * (Cdot + bmax(C, real(0)) * d.step.inv_dt);
      const real old = j->lower_impulse;
      j->lower_impulse = bmax(j->lower_impulse + impulse, real(0));
      impulse = j->lower_impulse - old;
      prismatic_push(j, impulse * j->axis, impulse * j->a1, impulse * j->a2, vA, wA, vB, wB);
    }
    {
      const real C = j->upper_translation - j->translation;
      const real Cdot = dot(j->axis, vA - vB) + j->a1 * wA - j->a2 * wB;
      real impulse = -j->axial_mass * (Cdot + bmax(C, real(0)) * d.step.inv_dt);
      const real old = j->upper_impulse;
      j->upper_impulse = bmax(j->upper_impulse + impulse, real(0));
      impulse = j->upper_impulse - old;
      const Vec2 P = impulse * j->axis;
      const real LA = impulse * j->a1, LB = impulse * j->a2;
      vA += j->inv_mass_a * P;
      wA += j->inv_ia * LA;
      vB -= j->inv_mass_b * P;
      wB -= j->inv_ib * LB;
    }
  }
  {
    const Vec2 Cdot(dot(j->perp, vB - vA) + j->s2 * wB - j->s1 * wA, wB - wA);
    const Vec2 df = mat22_solve(j->K, -Cdot);
    j->impulse2 += df;
    const Vec2 P = df.x * j->perp;
    const real LA = df.x * j->s1 + df.y;
    const real LB = df.x * j->s2 + df.y;
    prismatic_push(j, P, LA, LB, vA, wA, vB, wB);
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline bool prismatic_solve_position(Joint* j, const JointSolverData& d) {
  Vec2 cA = d.positions[j->index_a].c;
  real aA = d.positions[j->index_a].a;
  Vec2 cB = d.positions[j->index_b].c;
  real aB = d.positions[j->index_b].a;
  Rot qA, qB;
  qA.set(aA);
  qB.set(aB);
  const real mA = j->inv_mass_a, mB = j->inv_mass_b, iA = j->inv_ia, iB = j->inv_ib;
  const Vec2 rA = mul(qA, j->local_anchor_a - j->local_center_a);
  const Vec2 rB = mul(qB, j->local_anchor_b - j->local_center_b);
  const Vec2 dd = cB + rB - cA - rA;
  const Vec2 axis = mul(qA, j->local_xaxis_a);
  const real a1 = cross(dd + rA, axis);
  const real a2 = cross(rB, axis);
  const Vec2 perp = mul(qA, j->local_yaxis_a);
  const real s1 = cross(dd + rA, perp);
  const real s2 = cross(rB, perp);
  Vec3 impulse;
  const Vec2 C1(dot(perp, dd), aB - aA - j->reference_angle);
  real linear_error = babs(C1.x);
  const real angular_error = babs(C1.y);
  bool active = false;
  real C2 = 0;
  if (j->enable_limit) {
    const real translation = dot(axis, dd);
    if (babs(j->upper_translation - j->lower_translation) < real(2.0) * kLinearSlop) {
      C2 = translation;
      linear_error = bmax(linear_error, babs(translation));
      active = true;
    } else if (translation <= j->lower_translation) {
      C2 = bmin(translation - j->lower_translation, real(0));
      linear_error = bmax(linear_error, j->lower_translation - translation);
      active = true;
    } else if (translation >= j->upper_translation) {
      C2 = bmax(translation - j->upper_translation, real(0));
      linear_error = bmax(linear_error, translation - j->upper_translation);
      active = true;
    }
  }
  if (active) {
    const real k11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
    const real k12 = iA * s1 + iB * s2;
    const real k13 = iA * s1 * a1 + iB * s2 * a2;
    real k22 = iA + iB;
    if (k22 == 0) k22 = 1;
    const real k23 = iA * a1 + iB * a2;
    const real k33 = mA + mB + iA * a1 * a1 + iB * a2 * a2;
    Mat33 K;
    K.ex = Vec3(k11, k12, k13);
    K.ey = Vec3(k12, k22, k23);
    K.ez = Vec3(k13, k23, k33);
    impulse = solve33(K, Vec3(-C1.x, -C1.y, -C2));
  } else {
    const real k11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
    const real k12 = iA * s1 + iB * s2;
    real k22 = iA + iB;
    if (k22 == 0) k22 = 1;
    Mat22 K;
    K.ex = Vec2(k11, k12);
    K.ey = Vec2(k12, k22);
    const Vec2 impulse1 = mat22_solve(K, -C1);
    impulse = Vec3(impulse1.x, impulse1.y, 0);
  }
  const Vec2 P = impulse.x * perp + impulse.z * axis;
  const real LA = impulse.x * s1 + impulse.y + impulse.z * a1;
  const real LB = impulse.x * s2 + impulse.y + impulse.z * a2;
  cA -= mA * P;
  aA -= iA * LA;
  cB += mB * P;
  aB += iB * LB;
  d.positions[j->index_a].c = cA;
  d.positions[j->index_a].a = aA;
  d.positions[j->index_b].c = cB;
  d.positions[j->index_b].a = aB;
  return linear_error <= kLinearSlop && angular_error <= kAngularSlop;
}

inline void joint_init_velocity(Joint* j, const JointSolverData& d) {
  if (j->type == kDistanceJoint) distance_init_velocity(j, d);
  else if (j->type == kWeldJoint) weld_init_velocity(j, d);
  else if (j->type == kPrismaticJoint) prismatic_init_velocity(j, d);
  else revolute_init_velocity(j, d);
}
inline void joint_solve_velocity(Joint* j, const JointSolverData& d) {
  if (j->type == kDistanceJoint) distance_solve_velocity(j, d);
  else if (j->type == kWeldJoint) weld_solve_velocity(j, d);
  else if (j->type == kPrismaticJoint) prismatic_solve_velocity(j, d);
  else revolute_solve_velocity(j, d);
}
inline bool joint_solve_position(Joint* j, const JointSolverData& d) {
  if (j->type == kDistanceJoint) return distance_solve_position(j, d);
  if (j->type == kWeldJoint) return weld_solve_position(j, d);
  if (j->type == kPrismaticJoint) return prismatic_solve_position(j, d);
  return revolute_solve_position(j, d);
}

}  // namespace b2o
