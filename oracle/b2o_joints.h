// Oracle (test infrastructure, never part of the product): the distance and revolute joint
// solvers of the reference, restated in operation order.
//   distance joint  src/private/dynamics/joints/b2_distance_joint.rs:23-330
//   revolute joint  src/private/dynamics/joints/b2_revolute_joint.rs:60-330, construction
//                   src/joints/b2_revolute_joint.rs:253-287
//   B2Mat22::solve  src/b2_math.rs:278-293
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

inline void joint_init_velocity(Joint* j, const JointSolverData& d) {
  if (j->type == kDistanceJoint) distance_init_velocity(j, d);
  else revolute_init_velocity(j, d);
}
inline void joint_solve_velocity(Joint* j, const JointSolverData& d) {
  if (j->type == kDistanceJoint) distance_solve_velocity(j, d);
  else revolute_solve_velocity(j, d);
}
inline bool joint_solve_position(Joint* j, const JointSolverData& d) {
  return j->type == kDistanceJoint ? distance_solve_position(j, d) : revolute_solve_position(j, d);
}

}  // namespace b2o
