// Oracle (test infrastructure, never part of the product): the joint solvers of the reference,
// restated in operation order.
//   distance joint  src/private/dynamics/joints/b2_distance_joint.rs:23-330
//   revolute joint  src/private/dynamics/joints/b2_revolute_joint.rs:60-330, construction
//                   src/joints/b2_revolute_joint.rs:253-287
//   weld joint      src/private/dynamics/joints/b2_weld_joint.rs:23-300
//   prismatic joint src/private/dynamics/joints/b2_prismatic_joint.rs:40-460
//   friction joint  src/private/dynamics/joints/b2_friction_joint.rs:19-167
//   motor joint     src/private/dynamics/joints/b2_motor_joint.rs:18-198
//   mouse joint     src/private/dynamics/joints/b2_mouse_joint.rs:7-110
//   wheel joint     src/private/dynamics/joints/b2_wheel_joint.rs:7-400
//   pulley joint    src/private/dynamics/joints/b2_pulley_joint.rs:48-276
//   gear joint      src/private/dynamics/joints/b2_gear_joint.rs:8-436
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


// ---- friction and motor joints -------------------------------------------------------------
// src/private/dynamics/joints/b2_friction_joint.rs:19-167 and b2_motor_joint.rs:18-198: a point
// and an angle held by friction bounded by max force and torque; the motor joint drives body B to
// a linear and angular offset from body A (its anchors are the offset and B's centre), the same
// code with the position errors added to the velocity targets
inline void friction_motor_init_velocity(Joint* j, const JointSolverData& d) {
  joint_bind(j);
  const bool motor = j->type == kMotorJoint;
  const Vec2 cA = d.positions[j->index_a].c, cB = d.positions[j->index_b].c;
  const real aA = d.positions[j->index_a].a, aB = d.positions[j->index_b].a;
  Vec2 vA = d.velocities[j->index_a].v, vB = d.velocities[j->index_b].v;
  real wA = d.velocities[j->index_a].w, wB = d.velocities[j->index_b].w;
  Rot qA, qB;
  qA.set(aA);
  qB.set(aB);
  if (motor) {
    j->rA = mul(qA, j->linear_offset - j->local_center_a);
    j->rB = mul(qB, -j->local_center_b);
  } else {
    j->rA = mul(qA, j->local_anchor_a - j->local_center_a);
    j->rB = mul(qB, j->local_anchor_b - j->local_center_b);
  }
  const real mA = j->inv_mass_a, mB = j->inv_mass_b, iA = j->inv_ia, iB = j->inv_ib;
  Mat22 K;
  K.ex.x = mA + mB + iA * j->rA.y * j->rA.y + iB * j->rB.y * j->rB.y;
  K.ex.y = -iA * j->rA.x * j->rA.y - iB * j->rB.x * j->rB.y;
  K.ey.x = K.ex.y;
  K.ey.y = mA + mB + iA * j->rA.x * j->rA.x + iB * j->rB.x * j->rB.x;
  j->linear_mass = K.get_inverse();
  j->angular_mass = iA + iB;
  if (j->angular_mass > 0) j->angular_mass = real(1.0) / j->angular_mass;
  if (motor) {
    j->linear_error = cB + j->rB - cA - j->rA;
    j->angular_error = aB - aA - j->angular_offset;
  }
  if (d.step.warm_starting) {
    j->linear_impulse *= d.step.dt_ratio;
    j->angular_impulse *= d.step.dt_ratio;
    const Vec2 P = j->linear_impulse;
    vA -= mA * P;
    wA -= iA * (cross(j->rA, P) + j->angular_impulse);
    vB += mB * P;
    wB += iB * (cross(j->rB, P) + j->angular_impulse);
  } else {
    j->linear_impulse.set_zero();
    j->angular_impulse = 0;
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline void friction_motor_solve_velocity(Joint* j, const JointSolverData& d) {
  const bool motor = j->type == kMotorJoint;
  Vec2 vA = d.velocities[j->index_a].v, vB = d.velocities[j->index_b].v;
  real wA = d.velocities[j->index_a].w, wB = d.velocities[j->index_b].w;
  const real mA = j->inv_mass_a, mB = j->inv_mass_b, iA = j->inv_ia, iB = j->inv_ib;
  const real h = d.step.dt, inv_h = d.step.inv_dt;
  {
    real cdot = wB - wA;
    if (motor) cdot = cdot + inv_h * j->correction_factor * j->angular_error;
    real impulse = -j->angular_mass * cdot;
    const real old = j->angular_impulse;
    const real max_impulse = h * j->max_torque;
    j->angular_impulse = bclamp(j->angular_impulse + impulse, -max_impulse, max_impulse);
    impulse = j->angular_impulse - old;
    wA -= iA * impulse;
    wB += iB * impulse;
  }
  {
    Vec2 cdot = vB + cross(wB, j->rB) - vA - cross(wA, j->rA);
    if (motor) cdot = cdot + (inv_h * j->correction_factor) * j->linear_error;
    Vec2 impulse = -mul(j->linear_mass, cdot);
    const Vec2 old = j->linear_impulse;
    j->linear_impulse += impulse;
    const real max_impulse = h * j->max_force;
    if (j->linear_impulse.length_squared() > max_impulse * max_impulse) {
      j->linear_impulse.normalize();
      j->linear_impulse *= max_impulse;
    }
    impulse = j->linear_impulse - old;
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

// ---- mouse joint: src/private/dynamics/joints/b2_mouse_joint.rs:7-110 -----------------------
// a soft point constraint pulling body B's anchor to the target; body A takes no part
inline void mouse_init_velocity(Joint* j, const JointSolverData& d) {
  joint_bind(j);
  const Vec2 cB = d.positions[j->index_b].c;
  const real aB = d.positions[j->index_b].a;
  Vec2 vB = d.velocities[j->index_b].v;
  real wB = d.velocities[j->index_b].w;
  Rot qB;
  qB.set(aB);
  const real dmp = j->damping, k = j->stiffness, h = d.step.dt;
  j->gamma = h * (dmp + h * k);
  if (j->gamma != 0) j->gamma = real(1.0) / j->gamma;
  j->beta = h * k * j->gamma;
  j->rB = mul(qB, j->local_anchor_b - j->local_center_b);
  Mat22 K;
  K.ex.x = j->inv_mass_b + j->inv_ib * j->rB.y * j->rB.y + j->gamma;
  K.ex.y = -j->inv_ib * j->rB.x * j->rB.y;
  K.ey.x = K.ex.y;
  K.ey.y = j->inv_mass_b + j->inv_ib * j->rB.x * j->rB.x + j->gamma;
  j->mouse_mass = K.get_inverse();
  j->mouse_c = cB + j->rB - j->target;
  j->mouse_c *= j->beta;
  wB *= real(0.98);
  if (d.step.warm_starting) {
    j->impulse2 *= d.step.dt_ratio;
    vB += j->inv_mass_b * j->impulse2;
    wB += j->inv_ib * cross(j->rB, j->impulse2);
  } else {
    j->impulse2.set_zero();
  }
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline void mouse_solve_velocity(Joint* j, const JointSolverData& d) {
  Vec2 vB = d.velocities[j->index_b].v;
  real wB = d.velocities[j->index_b].w;
  const Vec2 cdot = vB + cross(wB, j->rB);
  Vec2 impulse = mul(j->mouse_mass, -(cdot + j->mouse_c + j->gamma * j->impulse2));
  const Vec2 old = j->impulse2;
  j->impulse2 += impulse;
  const real max_impulse = d.step.dt * j->max_force;
  if (j->impulse2.length_squared() > max_impulse * max_impulse) j->impulse2 *= max_impulse / j->impulse2.length();
  impulse = j->impulse2 - old;
  vB += j->inv_mass_b * impulse;
  wB += j->inv_ib * cross(j->rB, impulse);
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

// ---- wheel joint: src/private/dynamics/joints/b2_wheel_joint.rs:7-400 ------------------------
inline void wheel_init_velocity(Joint* j, const JointSolverData& d) {
  joint_bind(j);
  const real mA = j->inv_mass_a, mB = j->inv_mass_b, iA = j->inv_ia, iB = j->inv_ib;
  const Vec2 cA = d.positions[j->index_a].c, cB = d.positions[j->index_b].c;
  const real aA = d.positions[j->index_a].a, aB = d.positions[j->index_b].a;
  Vec2 vA = d.velocities[j->index_a].v, vB = d.velocities[j->index_b].v;
  real wA = d.velocities[j->index_a].w, wB = d.velocities[j->index_b].w;
  Rot qA, qB;
  qA.set(aA);
  qB.set(aB);
  const Vec2 rA = mul(qA, j->local_anchor_a - j->local_center_a);
  const Vec2 rB = mul(qB, j->local_anchor_b - j->local_center_b);
  const Vec2 dd = cB + rB - cA - rA;
  j->ay = mul(qA, j->local_yaxis_a);
  j->s_ay = cross(dd + rA, j->ay);
  j->s_by = cross(rB, j->ay);
  j->mass = mA + mB + iA * j->s_ay * j->s_ay + iB * j->s_by * j->s_by;
  if (j->mass > 0) j->mass = real(1.0) / j->mass;
  j->ax = mul(qA, j->local_xaxis_a);
  j->s_ax = cross(dd + rA, j->ax);
  j->s_bx = cross(rB, j->ax);
  const real inv_mass = mA + mB + iA * j->s_ax * j->s_ax + iB * j->s_bx * j->s_bx;
  j->axial_mass = inv_mass > 0 ? real(1.0) / inv_mass : real(0);
  j->spring_mass = 0;
  j->bias = 0;
  j->gamma = 0;
  if (j->stiffness > 0 && inv_mass > 0) {
    j->spring_mass = real(1.0) / inv_mass;
    const real C = dot(dd, j->ax);
    const real h = d.step.dt;
    j->gamma = h * (j->damping + h * j->stiffness);
    if (j->gamma > 0) j->gamma = real(1.0) / j->gamma;
    j->bias = C * h * j->stiffness * j->gamma;
    j->spring_mass = inv_mass + j->gamma;
    if (j->spring_mass > 0) j->spring_mass = real(1.0) / j->spring_mass;
  } else {
    j->spring_impulse = 0;
  }
  if (j->enable_limit) {
    j->translation = dot(j->ax, dd);
  } else {
    j->lower_impulse = 0;
    j->upper_impulse = 0;
  }
  if (j->enable_motor) {
    j->motor_mass = iA + iB;
    if (j->motor_mass > 0) j->motor_mass = real(1.0) / j->motor_mass;
  } else {
    j->motor_mass = 0;
    j->motor_impulse = 0;
  }
  if (d.step.warm_starting) {
    j->impulse *= d.step.dt_ratio;
    j->spring_impulse *= d.step.dt_ratio;
    j->motor_impulse *= d.step.dt_ratio;
    const real axial_impulse = j->spring_impulse + j->lower_impulse - j->upper_impulse;
    const Vec2 P = j->impulse * j->ay + axial_impulse * j->ax;
    const real LA = j->impulse * j->s_ay + axial_impulse * j->s_ax + j->motor_impulse;
    const real LB = j->impulse * j->s_by + axial_impulse * j->s_bx + j->motor_impulse;
    vA -= j->inv_mass_a * P;
    wA -= j->inv_ia * LA;
    vB += j->inv_mass_b * P;
    wB += j->inv_ib * LB;
  } else {
    j->impulse = 0;
    j->spring_impulse = 0;
    j->motor_impulse = 0;
    j->lower_impulse = 0;
    j->upper_impulse = 0;
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline void wheel_push(Joint* j, real impulse, Vec2 axis, real sa, real sb, Vec2& vA, real& wA, Vec2& vB, real& wB) {
  const Vec2 P = impulse * axis;
  const real LA = impulse * sa, LB = impulse * sb;
  vA -= j->inv_mass_a * P;
  wA -= j->inv_ia * LA;
  vB += j->inv_mass_b * P;
  wB += j->inv_ib * LB;
}

inline void wheel_solve_velocity(Joint* j, const JointSolverData& d) {
  const real mA = j->inv_mass_a, mB = j->inv_mass_b, iA = j->inv_ia, iB = j->inv_ib;
  Vec2 vA = d.velocities[j->index_a].v, vB = d.velocities[j->index_b].v;
  real wA = d.velocities[j->index_a].w, wB = d.velocities[j->index_b].w;
  (void)mA;
  (void)mB;
  {
    const real cdot = dot(j->ax, vB - vA) + j->s_bx * wB - j->s_ax * wA;
    const real impulse = -j->spring_mass * (cdot + j->bias + j->gamma * j->spring_impulse);
    j->spring_impulse += impulse;
    wheel_push(j, impulse, j->ax, j->s_ax, j->s_bx, vA, wA, vB, wB);
  }
  {
    const real cdot = wB - wA - j->motor_speed;
    real impulse = -j->motor_mass * cdot;
    const real old = j->motor_impulse;
    const real max_impulse = d.step.dt * j->max_motor_torque;
    j->motor_impulse = bclamp(j->motor_impulse + impulse, -max_impulse, max_impulse);
    impulse = j->motor_impulse - old;
    wA -= iA * impulse;
    wB += iB * impulse;
  }
  if (j->enable_limit) {
    {
      const real C = j->translation - j->lower_translation;
      const real cdot = dot(j->ax, vB - vA) + j->s_bx * wB - j->s_ax * wA;
      real impulse = -j->axial_mass * (cdot + bmax(C, real(0)) * d.step.inv_dt);
      const real old = j->lower_impulse;
      j->lower_impulse = bmax(j->lower_impulse + impulse, real(0));
      impulse = j->lower_impulse - old;
      wheel_push(j, impulse, j->ax, j->s_ax, j->s_bx, vA, wA, vB, wB);
    }
    {
      const real C = j->upper_translation - j->translation;
      const real cdot = dot(j->ax, vA - vB) + j->s_ax * wA - j->s_bx * wB;
      real impulse = -j->axial_mass * (cdot + bmax(C, real(0)) * d.step.inv_dt);
      const real old = j->upper_impulse;
      j->upper_impulse = bmax(j->upper_impulse + impulse, real(0));
      impulse = j->upper_impulse - old;
      const Vec2 P = impulse * j->ax;
      const real LA = impulse * j->s_ax, LB = impulse * j->s_bx;
      vA += j->inv_mass_a * P;
      wA += j->inv_ia * LA;
      vB -= j->inv_mass_b * P;
      wB -= j->inv_ib * LB;
    }
  }
  {
    const real cdot = dot(j->ay, vB - vA) + j->s_by * wB - j->s_ay * wA;
    const real impulse = -j->mass * cdot;
    j->impulse += impulse;
    wheel_push(j, impulse, j->ay, j->s_ay, j->s_by, vA, wA, vB, wB);
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

// the reference's position solve reads the velocity phase's m_ax in the limit's s_ax / s_bx and
// m_s_ay / m_s_by in the perpendicular mass (src/private/dynamics/joints/b2_wheel_joint.rs:300-370)
inline bool wheel_solve_position(Joint* j, const JointSolverData& d) {
  Vec2 cA = d.positions[j->index_a].c, cB = d.positions[j->index_b].c;
  real aA = d.positions[j->index_a].a, aB = d.positions[j->index_b].a;
  real linear_error = 0;
  if (j->enable_limit) {
    Rot qA, qB;
    qA.set(aA);
    qB.set(aB);
    const Vec2 rA = mul(qA, j->local_anchor_a - j->local_center_a);
    const Vec2 rB = mul(qB, j->local_anchor_b - j->local_center_b);
    const Vec2 dd = (cB - cA) + rB - rA;
    const Vec2 ax = mul(qA, j->local_xaxis_a);
    const real sAx = cross(dd + rA, j->ax);
    const real sBx = cross(rB, j->ax);
    real C = 0;
    const real translation = dot(ax, dd);
    if (babs(j->upper_translation - j->lower_translation) < real(2.0) * kLinearSlop) C = translation;
    else if (translation <= j->lower_translation) C = bmin(translation - j->lower_translation, real(0));
    else if (translation >= j->upper_translation) C = bmax(translation - j->upper_translation, real(0));
    if (C != 0) {
      const real inv_mass = j->inv_mass_a + j->inv_mass_b + j->inv_ia * sAx * sAx + j->inv_ib * sBx * sBx;
      real impulse = 0;
      if (inv_mass != 0) impulse = -C / inv_mass;
      const Vec2 P = impulse * ax;
      const real LA = impulse * sAx, LB = impulse * sBx;
      cA -= j->inv_mass_a * P;
      aA -= j->inv_ia * LA;
      cB += j->inv_mass_b * P;
      aB += j->inv_ib * LB;
      linear_error = babs(C);
    }
  }
  {
    Rot qA, qB;
    qA.set(aA);
    qB.set(aB);
    const Vec2 rA = mul(qA, j->local_anchor_a - j->local_center_a);
    const Vec2 rB = mul(qB, j->local_anchor_b - j->local_center_b);
    const Vec2 dd = (cB - cA) + rB - rA;
    const Vec2 ay = mul(qA, j->local_yaxis_a);
    const real sAy = cross(dd + rA, ay);
    const real sBy = cross(rB, ay);
    const real C = dot(dd, ay);
    const real inv_mass = j->inv_mass_a + j->inv_mass_b + j->inv_ia * j->s_ay * j->s_ay + j->inv_ib * j->s_by * j->s_by;
    real impulse = 0;
    if (inv_mass != 0) impulse = -C / inv_mass;
    const Vec2 P = impulse * ay;
    const real LA = impulse * sAy, LB = impulse * sBy;
    cA -= j->inv_mass_a * P;
    aA -= j->inv_ia * LA;
    cB += j->inv_mass_b * P;
    aB += j->inv_ib * LB;
    linear_error = bmax(linear_error, babs(C));
  }
  d.positions[j->index_a].c = cA;
  d.positions[j->index_a].a = aA;
  d.positions[j->index_b].c = cB;
  d.positions[j->index_b].a = aB;
  return linear_error <= kLinearSlop;
}

// ---- pulley joint: src/private/dynamics/joints/b2_pulley_joint.rs:48-276 --------------------
// length_a + ratio * length_b stays constant
inline void pulley_axes(Vec2* u, real* length) {
  *length = u->length();
  if (*length > real(10.0) * kLinearSlop) *u *= real(1.0) / *length;
  else u->set_zero();
}

inline void pulley_init_velocity(Joint* j, const JointSolverData& d) {
  joint_bind(j);
  const Vec2 cA = d.positions[j->index_a].c, cB = d.positions[j->index_b].c;
  const real aA = d.positions[j->index_a].a, aB = d.positions[j->index_b].a;
  Vec2 vA = d.velocities[j->index_a].v, vB = d.velocities[j->index_b].v;
  real wA = d.velocities[j->index_a].w, wB = d.velocities[j->index_b].w;
  Rot qA, qB;
  qA.set(aA);
  qB.set(aB);
  j->rA = mul(qA, j->local_anchor_a - j->local_center_a);
  j->rB = mul(qB, j->local_anchor_b - j->local_center_b);
  j->u_a = cA + j->rA - j->ground_anchor_a;
  j->u_b = cB + j->rB - j->ground_anchor_b;
  real la, lb;
  pulley_axes(&j->u_a, &la);
  pulley_axes(&j->u_b, &lb);
  const real ruA = cross(j->rA, j->u_a), ruB = cross(j->rB, j->u_b);
  const real mA = j->inv_mass_a + j->inv_ia * ruA * ruA;
  const real mB = j->inv_mass_b + j->inv_ib * ruB * ruB;
  j->mass = mA + j->ratio * j->ratio * mB;
  if (j->mass > 0) j->mass = real(1.0) / j->mass;
  if (d.step.warm_starting) {
    j->impulse *= d.step.dt_ratio;
    const Vec2 PA = -(j->impulse) * j->u_a;
    const Vec2 PB = (-j->ratio * j->impulse) * j->u_b;
    vA += j->inv_mass_a * PA;
    wA += j->inv_ia * cross(j->rA, PA);
    vB += j->inv_mass_b * PB;
    wB += j->inv_ib * cross(j->rB, PB);
  } else {
    j->impulse = 0;
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline void pulley_solve_velocity(Joint* j, const JointSolverData& d) {
  Vec2 vA = d.velocities[j->index_a].v, vB = d.velocities[j->index_b].v;
  real wA = d.velocities[j->index_a].w, wB = d.velocities[j->index_b].w;
  const Vec2 vpA = vA + cross(wA, j->rA);
  const Vec2 vpB = vB + cross(wB, j->rB);
  const real cdot = -dot(j->u_a, vpA) - j->ratio * dot(j->u_b, vpB);
  const real impulse = -j->mass * cdot;
  j->impulse += impulse;
  const Vec2 PA = -impulse * j->u_a;
  const Vec2 PB = -j->ratio * impulse * j->u_b;
  vA += j->inv_mass_a * PA;
  wA += j->inv_ia * cross(j->rA, PA);
  vB += j->inv_mass_b * PB;
  wB += j->inv_ib * cross(j->rB, PB);
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
}

inline bool pulley_solve_position(Joint* j, const JointSolverData& d) {
  Vec2 cA = d.positions[j->index_a].c, cB = d.positions[j->index_b].c;
  real aA = d.positions[j->index_a].a, aB = d.positions[j->index_b].a;
  Rot qA, qB;
  qA.set(aA);
  qB.set(aB);
  const Vec2 rA = mul(qA, j->local_anchor_a - j->local_center_a);
  const Vec2 rB = mul(qB, j->local_anchor_b - j->local_center_b);
  Vec2 uA = cA + rA - j->ground_anchor_a;
  Vec2 uB = cB + rB - j->ground_anchor_b;
  real la, lb;
  pulley_axes(&uA, &la);
  pulley_axes(&uB, &lb);
  const real ruA = cross(rA, uA), ruB = cross(rB, uB);
  const real mA = j->inv_mass_a + j->inv_ia * ruA * ruA;
  const real mB = j->inv_mass_b + j->inv_ib * ruB * ruB;
  real mass = mA + j->ratio * j->ratio * mB;
  if (mass > 0) mass = real(1.0) / mass;
  const real C = j->constant - la - j->ratio * lb;
  const real linear_error = babs(C);
  const real impulse = -mass * C;
  const Vec2 PA = -impulse * uA;
  const Vec2 PB = -j->ratio * impulse * uB;
  cA += j->inv_mass_a * PA;
  aA += j->inv_ia * cross(rA, PA);
  cB += j->inv_mass_b * PB;
  aB += j->inv_ib * cross(rB, PB);
  d.positions[j->index_a].c = cA;
  d.positions[j->index_a].a = aA;
  d.positions[j->index_b].c = cB;
  d.positions[j->index_b].a = aB;
  return linear_error < kLinearSlop;
}

// ---- gear joint: src/private/dynamics/joints/b2_gear_joint.rs:8-436 -------------------------
// coordinate_a + ratio * coordinate_b stays constant, the coordinates being the angles or
// translations of the two revolute or prismatic joints it couples (bodies C and D are the
// coupled joints' bodies A)
inline void gear_init_velocity(Joint* j, const JointSolverData& d) {
  j->index_a = j->a->island_index;
  j->index_b = j->b->island_index;
  j->index_c = j->c->island_index;
  j->index_d = j->d->island_index;
  j->lc_a = j->a->sweep.local_center;
  j->lc_b = j->b->sweep.local_center;
  j->lc_c = j->c->sweep.local_center;
  j->lc_d = j->d->sweep.local_center;
  j->m_a = j->a->inv_mass;
  j->m_b = j->b->inv_mass;
  j->m_c = j->c->inv_mass;
  j->m_d = j->d->inv_mass;
  j->i_a = j->a->inv_I;
  j->i_b = j->b->inv_I;
  j->i_c = j->c->inv_I;
  j->i_d = j->d->inv_I;
  const real aA = d.positions[j->index_a].a, aB = d.positions[j->index_b].a;
  const real aC = d.positions[j->index_c].a, aD = d.positions[j->index_d].a;
  Vec2 vA = d.velocities[j->index_a].v, vB = d.velocities[j->index_b].v;
  Vec2 vC = d.velocities[j->index_c].v, vD = d.velocities[j->index_d].v;
  real wA = d.velocities[j->index_a].w, wB = d.velocities[j->index_b].w;
  real wC = d.velocities[j->index_c].w, wD = d.velocities[j->index_d].w;
  Rot qA, qB, qC, qD;
  qA.set(aA);
  qB.set(aB);
  qC.set(aC);
  qD.set(aD);
  j->mass = 0;
  if (j->type_a == kRevoluteJoint) {
    j->jv_ac.set_zero();
    j->jw_a = 1;
    j->jw_c = 1;
    j->mass += j->i_a + j->i_c;
  } else {
    const Vec2 u = mul(qC, j->local_axis_c);
    const Vec2 rC = mul(qC, j->local_anchor_c - j->lc_c);
    const Vec2 rA = mul(qA, j->local_anchor_a - j->lc_a);
    j->jv_ac = u;
    j->jw_c = cross(rC, u);
    j->jw_a = cross(rA, u);
    j->mass += j->m_c + j->m_a + j->i_c * j->jw_c * j->jw_c + j->i_a * j->jw_a * j->jw_a;
  }
  if (j->type_b == kRevoluteJoint) {
    j->jv_bd.set_zero();
    j->jw_b = j->ratio;
    j->jw_d = j->ratio;
    j->mass += j->ratio * j->ratio * (j->i_b + j->i_d);
  } else {
    const Vec2 u = mul(qD, j->local_axis_d);
    const Vec2 rD = mul(qD, j->local_anchor_d - j->lc_d);
    const Vec2 rB = mul(qB, j->local_anchor_b - j->lc_b);
    j->jv_bd = j->ratio * u;
    j->jw_d = j->ratio * cross(rD, u);
    j->jw_b = j->ratio * cross(rB, u);
    j->mass += j->ratio * j->ratio * (j->m_d + j->m_b) + j->i_d * j->jw_d * j->jw_d + j->i_b * j->jw_b * j->jw_b;
  }
  j->mass = j->mass > 0 ? real(1.0) / j->mass : real(0);
  // (the reference does not scale this impulse by dt_ratio)
  if (d.step.warm_starting) {
    vA += (j->m_a * j->impulse) * j->jv_ac;
    wA += j->i_a * j->impulse * j->jw_a;
    vB += (j->m_b * j->impulse) * j->jv_bd;
    wB += j->i_b * j->impulse * j->jw_b;
    vC -= (j->m_c * j->impulse) * j->jv_ac;
    wC -= j->i_c * j->impulse * j->jw_c;
    vD -= (j->m_d * j->impulse) * j->jv_bd;
    wD -= j->i_d * j->impulse * j->jw_d;
  } else {
    j->impulse = 0;
  }
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
  d.velocities[j->index_c].v = vC;
  d.velocities[j->index_c].w = wC;
  d.velocities[j->index_d].v = vD;
  d.velocities[j->index_d].w = wD;
}

inline void gear_solve_velocity(Joint* j, const JointSolverData& d) {
  Vec2 vA = d.velocities[j->index_a].v, vB = d.velocities[j->index_b].v;
  Vec2 vC = d.velocities[j->index_c].v, vD = d.velocities[j->index_d].v;
  real wA = d.velocities[j->index_a].w, wB = d.velocities[j->index_b].w;
  real wC = d.velocities[j->index_c].w, wD = d.velocities[j->index_d].w;
  real cdot = dot(j->jv_ac, vA - vC) + dot(j->jv_bd, vB - vD);
  cdot += (j->jw_a * wA - j->jw_c * wC) + (j->jw_b * wB - j->jw_d * wD);
  const real impulse = -j->mass * cdot;
  j->impulse += impulse;
  vA += (j->m_a * impulse) * j->jv_ac;
  wA += j->i_a * impulse * j->jw_a;
  vB += (j->m_b * impulse) * j->jv_bd;
  wB += j->i_b * impulse * j->jw_b;
  vC -= (j->m_c * impulse) * j->jv_ac;
  wC -= j->i_c * impulse * j->jw_c;
  vD -= (j->m_d * impulse) * j->jv_bd;
  wD -= j->i_d * impulse * j->jw_d;
  d.velocities[j->index_a].v = vA;
  d.velocities[j->index_a].w = wA;
  d.velocities[j->index_b].v = vB;
  d.velocities[j->index_b].w = wB;
  d.velocities[j->index_c].v = vC;
  d.velocities[j->index_c].w = wC;
  d.velocities[j->index_d].v = vD;
  d.velocities[j->index_d].w = wD;
}

inline bool gear_solve_position(Joint* j, const JointSolverData& d) {
  Vec2 cA = d.positions[j->index_a].c, cB = d.positions[j->index_b].c;
  Vec2 cC = d.positions[j->index_c].c, cD = d.positions[j->index_d].c;
  real aA = d.positions[j->index_a].a, aB = d.positions[j->index_b].a;
  real aC = d.positions[j->index_c].a, aD = d.positions[j->index_d].a;
  Rot qA, qB, qC, qD;
  qA.set(aA);
  qB.set(aB);
  qC.set(aC);
  qD.set(aD);
  const real linear_error = 0;
  real coordinate_a, coordinate_b, jw_a, jw_b, jw_c, jw_d;
  Vec2 jv_ac, jv_bd;
  real mass = 0;
  if (j->type_a == kRevoluteJoint) {
    jv_ac.set_zero();
    jw_a = 1;
    jw_c = 1;
    mass += j->i_a + j->i_c;
    coordinate_a = aA - aC - j->reference_angle_a;
  } else {
    const Vec2 u = mul(qC, j->local_axis_c);
    const Vec2 rC = mul(qC, j->local_anchor_c - j->lc_c);
    const Vec2 rA = mul(qA, j->local_anchor_a - j->lc_a);
    jv_ac = u;
    jw_c = cross(rC, u);
    jw_a = cross(rA, u);
    mass += j->m_c + j->m_a + j->i_c * jw_c * jw_c + j->i_a * jw_a * jw_a;
    const Vec2 pC = j->local_anchor_c - j->lc_c;
    const Vec2 pA = mulT(qC, rA + (cA - cC));
    coordinate_a = dot(pA - pC, j->local_axis_c);
  }
  if (j->type_b == kRevoluteJoint) {
    jv_bd.set_zero();
    jw_b = j->ratio;
    jw_d = j->ratio;
    mass += j->ratio * j->ratio * (j->i_b + j->i_d);
    coordinate_b = aB - aD - j->reference_angle_b;
  } else {
    const Vec2 u = mul(qD, j->local_axis_d);
    const Vec2 rD = mul(qD, j->local_anchor_d - j->lc_d);
    const Vec2 rB = mul(qB, j->local_anchor_b - j->lc_b);
    jv_bd = j->ratio * u;
    jw_d = j->ratio * cross(rD, u);
    jw_b = j->ratio * cross(rB, u);
    mass += j->ratio * j->ratio * (j->m_d + j->m_b) + j->i_d * jw_d * jw_d + j->i_b * jw_b * jw_b;
    const Vec2 pD = j->local_anchor_d - j->lc_d;
    const Vec2 pB = mulT(qD, rB + (cB - cD));
    coordinate_b = dot(pB - pD, j->local_axis_d);
  }
  const real C = (coordinate_a + j->ratio * coordinate_b) - j->constant;
  real impulse = 0;
  if (mass > 0) impulse = -C / mass;
  cA += j->m_a * impulse * jv_ac;
  aA += j->i_a * impulse * jw_a;
  cB += j->m_b * impulse * jv_bd;
  aB += j->i_b * impulse * jw_b;
  cC -= j->m_c * impulse * jv_ac;
  aC -= j->i_c * impulse * jw_c;
  cD -= j->m_d * impulse * jv_bd;
  aD -= j->i_d * impulse * jw_d;
  d.positions[j->index_a].c = cA;
  d.positions[j->index_a].a = aA;
  d.positions[j->index_b].c = cB;
  d.positions[j->index_b].a = aB;
  d.positions[j->index_c].c = cC;
  d.positions[j->index_c].a = aC;
  d.positions[j->index_d].c = cD;
  d.positions[j->index_d].a = aD;
  return linear_error < kLinearSlop;
}

inline void joint_init_velocity(Joint* j, const JointSolverData& d) {
  switch (j->type) {
    case kDistanceJoint: distance_init_velocity(j, d); break;
    case kWeldJoint: weld_init_velocity(j, d); break;
    case kPrismaticJoint: prismatic_init_velocity(j, d); break;
    case kFrictionJoint:
    case kMotorJoint: friction_motor_init_velocity(j, d); break;
    case kMouseJoint: mouse_init_velocity(j, d); break;
    case kWheelJoint: wheel_init_velocity(j, d); break;
    case kPulleyJoint: pulley_init_velocity(j, d); break;
    case kGearJoint: gear_init_velocity(j, d); break;
    default: revolute_init_velocity(j, d); break;
  }
}
inline void joint_solve_velocity(Joint* j, const JointSolverData& d) {
  switch (j->type) {
    case kDistanceJoint: distance_solve_velocity(j, d); break;
    case kWeldJoint: weld_solve_velocity(j, d); break;
    case kPrismaticJoint: prismatic_solve_velocity(j, d); break;
    case kFrictionJoint:
    case kMotorJoint: friction_motor_solve_velocity(j, d); break;
    case kMouseJoint: mouse_solve_velocity(j, d); break;
    case kWheelJoint: wheel_solve_velocity(j, d); break;
    case kPulleyJoint: pulley_solve_velocity(j, d); break;
    case kGearJoint: gear_solve_velocity(j, d); break;
    default: revolute_solve_velocity(j, d); break;
  }
}
// friction, motor and mouse joints have no position constraint (they return true)
inline bool joint_solve_position(Joint* j, const JointSolverData& d) {
  switch (j->type) {
    case kDistanceJoint: return distance_solve_position(j, d);
    case kWeldJoint: return weld_solve_position(j, d);
    case kPrismaticJoint: return prismatic_solve_position(j, d);
    case kFrictionJoint:
    case kMotorJoint:
    case kMouseJoint: return true;
    case kWheelJoint: return wheel_solve_position(j, d);
    case kPulleyJoint: return pulley_solve_position(j, d);
    case kGearJoint: return gear_solve_position(j, d);
    default: return revolute_solve_position(j, d);
  }
}

}  // namespace b2o
