// b2g joints in the coloured island solve: distance, revolute, weld, prismatic, wheel, friction,
// motor, mouse, pulley and gear joints.
//
// The reference solves an island's joints in island order before its contacts in every velocity
// iteration, and after its contacts in every position iteration
// (src/private/dynamics/b2_island_private.rs:196-275). Here the joints of a world are coloured
// once, when they are created (no two joints of a colour share a dynamic body), and each colour is
// one launch: joint colours, then contact colours, per velocity iteration; contact colours, then
// joint colours, per position iteration. That is the sequence of one sequential order (joints by
// colour, then contacts by colour). The per-joint arithmetic follows the reference's operation
// order:
//   distance joint  src/private/dynamics/joints/b2_distance_joint.rs:80-320
//   revolute joint  src/private/dynamics/joints/b2_revolute_joint.rs:60-330
//   weld joint      src/private/dynamics/joints/b2_weld_joint.rs:23-300
//   prismatic joint src/private/dynamics/joints/b2_prismatic_joint.rs:40-460 (the perpendicular
//                   axis is (-axis.y, axis.x): bit for bit what rotating cross(1, local axis) gives)
//   B2Mat22::solve  src/b2_math.rs:278-293; B2Mat33 solve33 / solve22 / get_inverse22 /
//   get_sym_inverse33 src/private/common/b2_math.rs:5-71; b2_mul_mat33 / b2_mul22 src/b2_math.rs:601-608
#include "b2g_internal.h"

namespace b2g {

static constexpr int kBlock = 256;
constexpr float kAngularSlop = 2.0f / 180.0f * kPi;           // src/b2_common.rs:43
constexpr float kMaxAngularCorrection = 8.0f / 180.0f * kPi;  // src/b2_common.rs:64

__device__ inline V2 mat22_solve(float a11, float a21, float a12, float a22, V2 b) {
  float det = a11 * a22 - a12 * a21;
  if (det != 0.0f) det = 1.0f / det;
  return v2(det * (a22 * b.x - a12 * b.y), det * (a11 * b.y - a21 * b.x));
}

// ---- weld joint helpers: the symmetric 3x3 K and its inverses ------------------------------------
struct Sym33 {
  float m11, m12, m13, m22, m23, m33;  // columns ex = (m11, m12, m13), ey = (m12, m22, m23), ez = (m13, m23, m33)
};
struct F3 {
  float x, y, z;
};
__device__ inline float dot3(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ inline F3 cross3(F3 a, F3 b) {
  return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline Sym33 weld_k(float mA, float mB, float iA, float iB, V2 rA, V2 rB) {
  Sym33 k;
  k.m11 = ((mA + mB) + rA.y * rA.y * iA) + rB.y * rB.y * iB;
  k.m12 = -rA.y * rA.x * iA - rB.y * rB.x * iB;
  k.m13 = -rA.y * iA - rB.y * iB;
  k.m22 = ((mA + mB) + rA.x * rA.x * iA) + rB.x * rB.x * iB;
  k.m23 = rA.x * iA + rB.x * iB;
  k.m33 = iA + iB;
  return k;
}
__device__ inline Sym33 inverse22(const Sym33& k) {
  const float a = k.m11, b = k.m12, c = k.m12, d = k.m22;
  float det = a * d - b * c;
  if (det != 0.0f) det = 1.0f / det;
  return Sym33{det * d, -det * b, 0.0f, det * a, 0.0f, 0.0f};
}
__device__ inline Sym33 sym_inverse33(const Sym33& k) {
  const F3 ex{k.m11, k.m12, k.m13}, ey{k.m12, k.m22, k.m23}, ez{k.m13, k.m23, k.m33};
  float det = dot3(ex, cross3(ey, ez));
  if (det != 0.0f) det = 1.0f / det;
  const float a11 = k.m11, a12 = k.m12, a13 = k.m13, a22 = k.m22, a23 = k.m23, a33 = k.m33;
  return Sym33{det * (a22 * a33 - a23 * a23), det * (a13 * a23 - a12 * a33), det * (a12 * a23 - a13 * a22),
               det * (a11 * a33 - a13 * a13), det * (a13 * a12 - a11 * a23), det * (a11 * a22 - a12 * a12)};
}
__device__ inline F3 mul33(const Sym33& m, F3 v) {
  return F3{(v.x * m.m11 + v.y * m.m12) + v.z * m.m13, (v.x * m.m12 + v.y * m.m22) + v.z * m.m23,
            (v.x * m.m13 + v.y * m.m23) + v.z * m.m33};
}
__device__ inline V2 mul22(const Sym33& m, V2 v) { return v2(m.m11 * v.x + m.m12 * v.y, m.m12 * v.x + m.m22 * v.y); }
__device__ inline F3 solve33(const Sym33& k, F3 b) {
  const F3 ex{k.m11, k.m12, k.m13}, ey{k.m12, k.m22, k.m23}, ez{k.m13, k.m23, k.m33};
  float det = dot3(ex, cross3(ey, ez));
  if (det != 0.0f) det = 1.0f / det;
  return F3{det * dot3(b, cross3(ey, ez)), det * dot3(ex, cross3(b, ez)), det * dot3(ex, cross3(ey, b))};
}

// a joint's two bodies as the solver holds them
struct JointBodies {
  int a, b;
  bool dyn_a, dyn_b;
  float mA, iA, mB, iB;
  V2 lcA, lcB;
};

__device__ inline bool joint_bodies(int j, const JointArgs& a, JointBodies* jb, int* root) {
  const int2 b = a.jb[j];
  jb->a = b.x;
  jb->b = b.y;
  const int ta = a.btype[b.x], tb = a.btype[b.y];
  jb->dyn_a = ta == 2;
  jb->dyn_b = tb == 2;
  const int anchor = ta != 0 ? b.x : b.y;  // a body of the island (not static)
  if (a.iv.asleep(anchor)) return false;
  *root = a.iv.parent[anchor];
  const float4 ma = a.mass[b.x], mb = a.mass[b.y];
  jb->mA = ma.x;
  jb->iA = ma.y;
  jb->lcA = v2(ma.z, ma.w);
  jb->mB = mb.x;
  jb->iB = mb.y;
  jb->lcB = v2(mb.z, mb.w);
  return true;
}

__device__ inline void prismatic_push(const JointBodies& B, V2 P, float LA, float LB, V2& vA, float& wA, V2& vB,
                                      float& wB) {
  vA = vA - B.mA * P;
  wA -= B.iA * LA;
  vB = vB + B.mB * P;
  wB += B.iB * LB;
}


// ---- friction, motor, mouse, wheel, pulley and gear joints --------------------------------------
//   friction joint  src/private/dynamics/joints/b2_friction_joint.rs:19-167
//   motor joint     src/private/dynamics/joints/b2_motor_joint.rs:18-198 (the friction joint's code with
//                   the anchors at the linear offset and B's centre and the position errors added)
//   mouse joint     src/private/dynamics/joints/b2_mouse_joint.rs:7-110 (body B only)
//   wheel joint     src/private/dynamics/joints/b2_wheel_joint.rs:7-400
//   pulley joint    src/private/dynamics/joints/b2_pulley_joint.rs:48-276
//   gear joint      src/private/dynamics/joints/b2_gear_joint.rs:140-436 (four bodies: C and D are the
//                   coupled joints' bodies A; no dt_ratio on its warm start, as the reference)
// Friction, motor and mouse joints have no position constraint.
struct BodyVel {
  V2 v;
  float w;
};
struct Mat22f {
  V2 ex, ey;  // columns (B2Mat22)
};
__device__ inline BodyVel load_vel(const JointArgs& a, int b) {
  const float4 q = a.vel[b];
  return BodyVel{v2(q.x, q.y), q.z};
}
__device__ inline void store_vel(const JointArgs& a, int b, bool dyn, const BodyVel& x) {
  if (dyn) a.vel[b] = make_float4(x.v.x, x.v.y, x.w, 0.0f);
}
__device__ inline Mat22f inverse22f(float exx, float exy, float eyx, float eyy) {
  float det = exx * eyy - eyx * exy;
  if (det != 0.0f) det = 1.0f / det;
  Mat22f m;
  m.ex = v2(det * eyy, -det * exy);
  m.ey = v2(-det * eyx, det * exx);
  return m;
}
__device__ inline V2 mul22f(const Mat22f& m, V2 v) { return v2(m.ex.x * v.x + m.ey.x * v.y, m.ex.y * v.x + m.ey.y * v.y); }

__device__ inline void pulley_axis(V2* u, float* length) {
  *length = len(*u);
  if (*length > 10.0f * kLinearSlop) *u = (1.0f / *length) * *u;
  else *u = v2(0.0f, 0.0f);
}

// the gear's other two bodies
struct GearBodies {
  int c, d;
  bool dyn_c, dyn_d;
  float mC, iC, mD, iD;
  V2 lcC, lcD;
};
__device__ inline GearBodies gear_bodies(int j, const JointArgs& a) {
  const int4 t = a.jt[j];
  GearBodies g;
  g.c = t.z;
  g.d = t.w;
  g.dyn_c = a.btype[g.c] == 2;
  g.dyn_d = a.btype[g.d] == 2;
  const float4 mc = a.mass[g.c], md = a.mass[g.d];
  g.mC = mc.x;
  g.iC = mc.y;
  g.lcC = v2(mc.z, mc.w);
  g.mD = md.x;
  g.iD = md.y;
  g.lcD = v2(md.z, md.w);
  return g;
}

__device__ void extra_init(int j, const JointArgs& a, const JointBodies& B, int type) {
  const float4 d0 = a.jd0[j], d1 = a.jd1[j], d2 = a.jd2[j];
  const float4 PA = a.pos[B.a], PB = a.pos[B.b];
  BodyVel A = load_vel(a, B.a), Bv = load_vel(a, B.b);
  const V2 cA = v2(PA.x, PA.y), cB = v2(PB.x, PB.y);
  const float aA = PA.z, aB = PB.z;
  const float mA = B.mA, mB = B.mB, iA = B.iA, iB = B.iB;
  float4 imp = a.jimp[j];
  if (type == B2G_JOINT_FRICTION || type == B2G_JOINT_MOTOR) {
    const bool motor = type == B2G_JOINT_MOTOR;
    const Rot qA = rot(aA), qB = rot(aB);
    V2 rA, rB;
    if (motor) {
      rA = mul(qA, v2(d0.x, d0.y) - B.lcA);
      rB = mul(qB, -B.lcB);
    } else {
      rA = mul(qA, v2(d0.x, d0.y) - B.lcA);
      rB = mul(qB, v2(d0.z, d0.w) - B.lcB);
    }
    const float kxx = ((mA + mB) + (iA * rA.y) * rA.y) + (iB * rB.y) * rB.y;
    const float kxy = ((-iA) * rA.x) * rA.y - (iB * rB.x) * rB.y;
    const float kyy = ((mA + mB) + (iA * rA.x) * rA.x) + (iB * rB.x) * rB.x;
    const Mat22f M = inverse22f(kxx, kxy, kxy, kyy);
    float angular_mass = iA + iB;
    if (angular_mass > 0.0f) angular_mass = 1.0f / angular_mass;
    V2 lin_err = v2(0.0f, 0.0f);
    float ang_err = 0.0f;
    if (motor) {
      lin_err = ((cB + rB) - cA) - rA;
      ang_err = (aB - aA) - d1.z;
    }
    if (a.warm) {
      imp.x *= a.dt_ratio;
      imp.y *= a.dt_ratio;
      imp.z *= a.dt_ratio;
      const V2 P = v2(imp.x, imp.y);
      A.v = A.v - mA * P;
      A.w -= iA * (cross(rA, P) + imp.z);
      Bv.v = Bv.v + mB * P;
      Bv.w += iB * (cross(rB, P) + imp.z);
    } else {
      imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    a.jr[j] = make_float4(rA.x, rA.y, rB.x, rB.y);
    a.jk[j] = make_float4(M.ex.x, M.ex.y, M.ey.x, M.ey.y);
    a.jk2[j] = make_float4(angular_mass, lin_err.x, lin_err.y, ang_err);
  } else if (type == B2G_JOINT_MOUSE) {
    const Rot qB = rot(aB);
    const float dmp = d2.x, k = d1.w, h = a.dt;
    float gamma = h * (dmp + h * k);
    if (gamma != 0.0f) gamma = 1.0f / gamma;
    const float beta = (h * k) * gamma;
    const V2 rB = mul(qB, v2(d0.z, d0.w) - B.lcB);
    const float kxx = (mB + (iB * rB.y) * rB.y) + gamma;
    const float kxy = ((-iB) * rB.x) * rB.y;
    const float kyy = (mB + (iB * rB.x) * rB.x) + gamma;
    const Mat22f M = inverse22f(kxx, kxy, kxy, kyy);
    V2 C = ((cB + rB) - v2(d1.x, d1.y));
    C = v2(C.x * beta, C.y * beta);
    Bv.w *= 0.98f;
    if (a.warm) {
      imp.x *= a.dt_ratio;
      imp.y *= a.dt_ratio;
      Bv.v = Bv.v + mB * v2(imp.x, imp.y);
      Bv.w += iB * cross(rB, v2(imp.x, imp.y));
    } else {
      imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    a.jr[j] = make_float4(rB.x, rB.y, C.x, C.y);
    a.jk[j] = make_float4(M.ex.x, M.ex.y, M.ey.x, M.ey.y);
    a.jk2[j] = make_float4(gamma, 0.0f, 0.0f, 0.0f);
    a.jimp[j] = imp;
    store_vel(a, B.b, B.dyn_b, Bv);
    return;
  } else if (type == B2G_JOINT_WHEEL) {
    const int flags = a.jt[j].y;
    const bool limit = flags & 2, motor = flags & 4;
    const float stiffness = d2.x, damping = d2.y;
    const V2 xaxis = v2(d2.z, d2.w), yaxis = cross(1.0f, xaxis);
    float upper = a.jimp2[j];
    const Rot qA = rot(aA), qB = rot(aB);
    const V2 rA = mul(qA, v2(d0.x, d0.y) - B.lcA);
    const V2 rB = mul(qB, v2(d0.z, d0.w) - B.lcB);
    const V2 dd = ((cB + rB) - cA) - rA;
    const V2 ay = mul(qA, yaxis);
    const float s_ay = cross(dd + rA, ay), s_by = cross(rB, ay);
    float mass = ((mA + mB) + (iA * s_ay) * s_ay) + (iB * s_by) * s_by;
    if (mass > 0.0f) mass = 1.0f / mass;
    const V2 ax = mul(qA, xaxis);
    const float s_ax = cross(dd + rA, ax), s_bx = cross(rB, ax);
    const float inv_mass = ((mA + mB) + (iA * s_ax) * s_ax) + (iB * s_bx) * s_bx;
    const float axial_mass = inv_mass > 0.0f ? 1.0f / inv_mass : 0.0f;
    float spring_mass = 0.0f, bias = 0.0f, gamma = 0.0f;
    if (stiffness > 0.0f && inv_mass > 0.0f) {
      const float C = dot(dd, ax);
      const float h = a.dt;
      gamma = h * (damping + h * stiffness);
      if (gamma > 0.0f) gamma = 1.0f / gamma;
      bias = ((C * h) * stiffness) * gamma;
      spring_mass = inv_mass + gamma;
      if (spring_mass > 0.0f) spring_mass = 1.0f / spring_mass;
    } else {
      imp.y = 0.0f;
    }
    float translation = 0.0f;
    if (limit) {
      translation = dot(ax, dd);
    } else {
      imp.w = 0.0f;
      upper = 0.0f;
    }
    float motor_mass = 0.0f;
    if (motor) {
      motor_mass = iA + iB;
      if (motor_mass > 0.0f) motor_mass = 1.0f / motor_mass;
    } else {
      imp.z = 0.0f;
    }
    if (a.warm) {
      imp.x *= a.dt_ratio;
      imp.y *= a.dt_ratio;
      imp.z *= a.dt_ratio;
      const float axial = (imp.y + imp.w) - upper;
      const V2 P = imp.x * ay + axial * ax;
      const float LA = (imp.x * s_ay + axial * s_ax) + imp.z;
      const float LB = (imp.x * s_by + axial * s_bx) + imp.z;
      A.v = A.v - mA * P;
      A.w -= iA * LA;
      Bv.v = Bv.v + mB * P;
      Bv.w += iB * LB;
    } else {
      imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      upper = 0.0f;
    }
    a.jimp2[j] = upper;
    a.jr[j] = make_float4(ax.x, ax.y, ay.x, ay.y);
    a.jk[j] = make_float4(s_ax, s_bx, s_ay, s_by);
    a.jk2[j] = make_float4(mass, motor_mass, axial_mass, spring_mass);
    a.jk3[j] = make_float4(bias, gamma, translation, 0.0f);
  } else if (type == B2G_JOINT_PULLEY) {
    const float ratio = d2.x;
    const V2 gA = v2(d1.x, d1.y), gB = v2(d1.z, d1.w);
    const Rot qA = rot(aA), qB = rot(aB);
    const V2 rA = mul(qA, v2(d0.x, d0.y) - B.lcA);
    const V2 rB = mul(qB, v2(d0.z, d0.w) - B.lcB);
    V2 uA = (cA + rA) - gA, uB = (cB + rB) - gB;
    float la, lb;
    pulley_axis(&uA, &la);
    pulley_axis(&uB, &lb);
    const float ruA = cross(rA, uA), ruB = cross(rB, uB);
    const float pmA = mA + (iA * ruA) * ruA, pmB = mB + (iB * ruB) * ruB;
    float mass = pmA + (ratio * ratio) * pmB;
    if (mass > 0.0f) mass = 1.0f / mass;
    if (a.warm) {
      imp.x *= a.dt_ratio;
      const V2 P1 = (-imp.x) * uA;
      const V2 P2 = ((-ratio) * imp.x) * uB;
      A.v = A.v + mA * P1;
      A.w += iA * cross(rA, P1);
      Bv.v = Bv.v + mB * P2;
      Bv.w += iB * cross(rB, P2);
    } else {
      imp.x = 0.0f;
    }
    a.jr[j] = make_float4(rA.x, rA.y, rB.x, rB.y);
    a.jk[j] = make_float4(uA.x, uA.y, uB.x, uB.y);
    a.jk2[j] = make_float4(mass, 0.0f, 0.0f, 0.0f);
  } else {  // gear
    const GearBodies G = gear_bodies(j, a);
    const float4 d3 = a.jd3[j];
    const int flags = a.jt[j].y;
    const float ratio = d3.z;
    const float4 PC = a.pos[G.c], PD = a.pos[G.d];
    BodyVel Cv = load_vel(a, G.c), Dv = load_vel(a, G.d);
    const Rot qA = rot(aA), qB = rot(aB), qC = rot(PC.z), qD = rot(PD.z);
    float mass = 0.0f;
    V2 jv_ac, jv_bd;
    float jw_a, jw_b, jw_c, jw_d;
    if (!(flags & 8)) {
      jv_ac = v2(0.0f, 0.0f);
      jw_a = 1.0f;
      jw_c = 1.0f;
      mass += iA + G.iC;
    } else {
      const V2 u = mul(qC, v2(d2.x, d2.y));
      const V2 rC = mul(qC, v2(d1.x, d1.y) - G.lcC);
      const V2 rA = mul(qA, v2(d0.x, d0.y) - B.lcA);
      jv_ac = u;
      jw_c = cross(rC, u);
      jw_a = cross(rA, u);
      mass += ((G.mC + mA) + (G.iC * jw_c) * jw_c) + (iA * jw_a) * jw_a;
    }
    if (!(flags & 16)) {
      jv_bd = v2(0.0f, 0.0f);
      jw_b = ratio;
      jw_d = ratio;
      mass += (ratio * ratio) * (iB + G.iD);
    } else {
      const V2 u = mul(qD, v2(d2.z, d2.w));
      const V2 rD = mul(qD, v2(d1.z, d1.w) - G.lcD);
      const V2 rB = mul(qB, v2(d0.z, d0.w) - B.lcB);
      jv_bd = ratio * u;
      jw_d = ratio * cross(rD, u);
      jw_b = ratio * cross(rB, u);
      mass += (((ratio * ratio) * (G.mD + mB)) + (G.iD * jw_d) * jw_d) + (iB * jw_b) * jw_b;
    }
    mass = mass > 0.0f ? 1.0f / mass : 0.0f;
    if (a.warm) {
      const float im = imp.x;
      A.v = A.v + (mA * im) * jv_ac;
      A.w += (iA * im) * jw_a;
      Bv.v = Bv.v + (mB * im) * jv_bd;
      Bv.w += (iB * im) * jw_b;
      Cv.v = Cv.v - (G.mC * im) * jv_ac;
      Cv.w -= (G.iC * im) * jw_c;
      Dv.v = Dv.v - (G.mD * im) * jv_bd;
      Dv.w -= (G.iD * im) * jw_d;
    } else {
      imp.x = 0.0f;
    }
    a.jr[j] = make_float4(jv_ac.x, jv_ac.y, jv_bd.x, jv_bd.y);
    a.jk[j] = make_float4(jw_a, jw_b, jw_c, jw_d);
    a.jk2[j] = make_float4(mass, 0.0f, 0.0f, 0.0f);
    store_vel(a, G.c, G.dyn_c, Cv);
    store_vel(a, G.d, G.dyn_d, Dv);
  }
  a.jimp[j] = imp;
  store_vel(a, B.a, B.dyn_a, A);
  store_vel(a, B.b, B.dyn_b, Bv);
}

__device__ inline void wheel_push(float impulse, V2 axis, float sa, float sb, const JointBodies& B, BodyVel& A,
                                  BodyVel& Bv) {
  const V2 P = impulse * axis;
  A.v = A.v - B.mA * P;
  A.w -= B.iA * (impulse * sa);
  Bv.v = Bv.v + B.mB * P;
  Bv.w += B.iB * (impulse * sb);
}

__device__ void extra_velocity(int j, const JointArgs& a, const JointBodies& B, int type) {
  BodyVel A = load_vel(a, B.a), Bv = load_vel(a, B.b);
  const float4 r = a.jr[j], k = a.jk[j], k2 = a.jk2[j];
  const float mA = B.mA, mB = B.mB, iA = B.iA, iB = B.iB;
  float4 imp = a.jimp[j];
  if (type == B2G_JOINT_FRICTION || type == B2G_JOINT_MOTOR) {
    const bool motor = type == B2G_JOINT_MOTOR;
    const float4 d1 = a.jd1[j];
    const float max_force = d1.x, max_torque = d1.y, cf = d1.w;
    const V2 rA = v2(r.x, r.y), rB = v2(r.z, r.w);
    const float h = a.dt, inv_h = a.inv_dt;
    {
      float cdot = Bv.w - A.w;
      if (motor) cdot = cdot + (inv_h * cf) * k2.w;
      float impulse = -k2.x * cdot;
      const float old = imp.z;
      const float max_impulse = h * max_torque;
      imp.z = bclamp(imp.z + impulse, -max_impulse, max_impulse);
      impulse = imp.z - old;
      A.w -= iA * impulse;
      Bv.w += iB * impulse;
    }
    {
      V2 cdot = ((Bv.v + cross(Bv.w, rB)) - A.v) - cross(A.w, rA);
      if (motor) cdot = cdot + (inv_h * cf) * v2(k2.y, k2.z);
      Mat22f M;
      M.ex = v2(k.x, k.y);
      M.ey = v2(k.z, k.w);
      V2 impulse = -mul22f(M, cdot);
      const V2 old = v2(imp.x, imp.y);
      V2 acc = old + impulse;
      const float max_impulse = h * max_force;
      if (dot(acc, acc) > max_impulse * max_impulse) {
        normalize(&acc);
        acc = v2(acc.x * max_impulse, acc.y * max_impulse);
      }
      imp.x = acc.x;
      imp.y = acc.y;
      impulse = acc - old;
      A.v = A.v - mA * impulse;
      A.w -= iA * cross(rA, impulse);
      Bv.v = Bv.v + mB * impulse;
      Bv.w += iB * cross(rB, impulse);
    }
  } else if (type == B2G_JOINT_MOUSE) {
    const float max_force = a.jd1[j].z;
    const V2 rB = v2(r.x, r.y), C = v2(r.z, r.w);
    const float gamma = k2.x;
    const V2 cdot = Bv.v + cross(Bv.w, rB);
    Mat22f M;
    M.ex = v2(k.x, k.y);
    M.ey = v2(k.z, k.w);
    const V2 old = v2(imp.x, imp.y);
    V2 impulse = mul22f(M, -((cdot + C) + gamma * old));
    V2 acc = old + impulse;
    const float max_impulse = a.dt * max_force;
    if (dot(acc, acc) > max_impulse * max_impulse) {
      const float s = max_impulse / len(acc);
      acc = v2(acc.x * s, acc.y * s);
    }
    imp.x = acc.x;
    imp.y = acc.y;
    impulse = acc - old;
    Bv.v = Bv.v + mB * impulse;
    Bv.w += iB * cross(rB, impulse);
    a.jimp[j] = imp;
    store_vel(a, B.b, B.dyn_b, Bv);
    return;
  } else if (type == B2G_JOINT_WHEEL) {
    const int flags = a.jt[j].y;
    const bool limit = flags & 2;
    const float4 d1 = a.jd1[j], k3 = a.jk3[j];
    const float lower = d1.x, upper_t = d1.y, max_torque = d1.z, motor_speed = d1.w;
    const V2 ax = v2(r.x, r.y), ay = v2(r.z, r.w);
    const float s_ax = k.x, s_bx = k.y, s_ay = k.z, s_by = k.w;
    const float mass = k2.x, motor_mass = k2.y, axial_mass = k2.z, spring_mass = k2.w;
    const float bias = k3.x, gamma = k3.y, translation = k3.z;
    float upper = a.jimp2[j];
    {
      const float cdot = (dot(ax, Bv.v - A.v) + s_bx * Bv.w) - s_ax * A.w;
      const float impulse = -spring_mass * ((cdot + bias) + gamma * imp.y);
      imp.y += impulse;
      wheel_push(impulse, ax, s_ax, s_bx, B, A, Bv);
    }
    {
      const float cdot = (Bv.w - A.w) - motor_speed;
      float impulse = -motor_mass * cdot;
      const float old = imp.z;
      const float max_impulse = a.dt * max_torque;
      imp.z = bclamp(imp.z + impulse, -max_impulse, max_impulse);
      impulse = imp.z - old;
      A.w -= iA * impulse;
      Bv.w += iB * impulse;
    }
    if (limit) {
      {
        const float C = translation - lower;
        const float cdot = (dot(ax, Bv.v - A.v) + s_bx * Bv.w) - s_ax * A.w;
        float impulse = -axial_mass * (cdot + bmax(C, 0.0f) * a.inv_dt);
        const float old = imp.w;
        imp.w = bmax(imp.w + impulse, 0.0f);
        impulse = imp.w - old;
        wheel_push(impulse, ax, s_ax, s_bx, B, A, Bv);
      }
      {
        const float C = upper_t - translation;
        const float cdot = (dot(ax, A.v - Bv.v) + s_ax * A.w) - s_bx * Bv.w;
        float impulse = -axial_mass * (cdot + bmax(C, 0.0f) * a.inv_dt);
        const float old = upper;
        upper = bmax(upper + impulse, 0.0f);
        impulse = upper - old;
        const V2 P = impulse * ax;
        A.v = A.v + mA * P;
        A.w += iA * (impulse * s_ax);
        Bv.v = Bv.v - mB * P;
        Bv.w -= iB * (impulse * s_bx);
      }
    }
    {
      const float cdot = (dot(ay, Bv.v - A.v) + s_by * Bv.w) - s_ay * A.w;
      const float impulse = -mass * cdot;
      imp.x += impulse;
      wheel_push(impulse, ay, s_ay, s_by, B, A, Bv);
    }
    a.jimp2[j] = upper;
  } else if (type == B2G_JOINT_PULLEY) {
    const float ratio = a.jd2[j].x;
    const V2 rA = v2(r.x, r.y), rB = v2(r.z, r.w), uA = v2(k.x, k.y), uB = v2(k.z, k.w);
    const V2 vpA = A.v + cross(A.w, rA), vpB = Bv.v + cross(Bv.w, rB);
    const float cdot = -dot(uA, vpA) - ratio * dot(uB, vpB);
    const float impulse = -k2.x * cdot;
    imp.x += impulse;
    const V2 P1 = (-impulse) * uA;
    const V2 P2 = ((-ratio) * impulse) * uB;
    A.v = A.v + mA * P1;
    A.w += iA * cross(rA, P1);
    Bv.v = Bv.v + mB * P2;
    Bv.w += iB * cross(rB, P2);
  } else {  // gear
    const GearBodies G = gear_bodies(j, a);
    BodyVel Cv = load_vel(a, G.c), Dv = load_vel(a, G.d);
    const V2 jv_ac = v2(r.x, r.y), jv_bd = v2(r.z, r.w);
    const float jw_a = k.x, jw_b = k.y, jw_c = k.z, jw_d = k.w;
    float cdot = dot(jv_ac, A.v - Cv.v) + dot(jv_bd, Bv.v - Dv.v);
    cdot += (jw_a * A.w - jw_c * Cv.w) + (jw_b * Bv.w - jw_d * Dv.w);
    const float impulse = -k2.x * cdot;
    imp.x += impulse;
    A.v = A.v + (mA * impulse) * jv_ac;
    A.w += (iA * impulse) * jw_a;
    Bv.v = Bv.v + (mB * impulse) * jv_bd;
    Bv.w += (iB * impulse) * jw_b;
    Cv.v = Cv.v - (G.mC * impulse) * jv_ac;
    Cv.w -= (G.iC * impulse) * jw_c;
    Dv.v = Dv.v - (G.mD * impulse) * jv_bd;
    Dv.w -= (G.iD * impulse) * jw_d;
    store_vel(a, G.c, G.dyn_c, Cv);
    store_vel(a, G.d, G.dyn_d, Dv);
  }
  a.jimp[j] = imp;
  store_vel(a, B.a, B.dyn_a, A);
  store_vel(a, B.b, B.dyn_b, Bv);
}

// returns whether the joint is within tolerance
__device__ bool extra_position(int j, const JointArgs& a, const JointBodies& B, int type) {
  if (type == B2G_JOINT_FRICTION || type == B2G_JOINT_MOTOR || type == B2G_JOINT_MOUSE) return true;
  const float4 d0 = a.jd0[j], d1 = a.jd1[j], d2 = a.jd2[j];
  const float4 PA = a.pos[B.a], PB = a.pos[B.b];
  V2 cA = v2(PA.x, PA.y), cB = v2(PB.x, PB.y);
  float aA = PA.z, aB = PB.z;
  const float mA = B.mA, mB = B.mB, iA = B.iA, iB = B.iB;
  const V2 laA = v2(d0.x, d0.y), laB = v2(d0.z, d0.w);
  bool okay;
  if (type == B2G_JOINT_WHEEL) {
    const bool limit = a.jt[j].y & 2;
    const float lower = d1.x, upper = d1.y;
    const V2 xaxis = v2(d2.z, d2.w), yaxis = cross(1.0f, xaxis);
    const float4 r = a.jr[j], k = a.jk[j];
    const V2 m_ax = v2(r.x, r.y);
    const float m_s_ay = k.z, m_s_by = k.w;
    float linear_error = 0.0f;
    if (limit) {
      const Rot qA = rot(aA), qB = rot(aB);
      const V2 rA = mul(qA, laA - B.lcA);
      const V2 rB = mul(qB, laB - B.lcB);
      const V2 dd = ((cB - cA) + rB) - rA;
      const V2 ax = mul(qA, xaxis);
      const float sAx = cross(dd + rA, m_ax), sBx = cross(rB, m_ax);
      float C = 0.0f;
      const float translation = dot(ax, dd);
      if (fabsf(upper - lower) < 2.0f * kLinearSlop) C = translation;
      else if (translation <= lower) C = bmin(translation - lower, 0.0f);
      else if (translation >= upper) C = bmax(translation - upper, 0.0f);
      if (C != 0.0f) {
        const float inv_mass = ((mA + mB) + (iA * sAx) * sAx) + (iB * sBx) * sBx;
        float impulse = 0.0f;
        if (inv_mass != 0.0f) impulse = -C / inv_mass;
        const V2 P = impulse * ax;
        cA = cA - mA * P;
        aA -= iA * (impulse * sAx);
        cB = cB + mB * P;
        aB += iB * (impulse * sBx);
        linear_error = fabsf(C);
      }
    }
    {
      const Rot qA = rot(aA), qB = rot(aB);
      const V2 rA = mul(qA, laA - B.lcA);
      const V2 rB = mul(qB, laB - B.lcB);
      const V2 dd = ((cB - cA) + rB) - rA;
      const V2 ay = mul(qA, yaxis);
      const float sAy = cross(dd + rA, ay), sBy = cross(rB, ay);
      const float C = dot(dd, ay);
      const float inv_mass = ((mA + mB) + (iA * m_s_ay) * m_s_ay) + (iB * m_s_by) * m_s_by;
      float impulse = 0.0f;
      if (inv_mass != 0.0f) impulse = -C / inv_mass;
      const V2 P = impulse * ay;
      cA = cA - mA * P;
      aA -= iA * (impulse * sAy);
      cB = cB + mB * P;
      aB += iB * (impulse * sBy);
      linear_error = bmax(linear_error, fabsf(C));
    }
    okay = linear_error <= kLinearSlop;
  } else if (type == B2G_JOINT_PULLEY) {
    const float ratio = d2.x, constant = d2.y;
    const V2 gA = v2(d1.x, d1.y), gB = v2(d1.z, d1.w);
    const Rot qA = rot(aA), qB = rot(aB);
    const V2 rA = mul(qA, laA - B.lcA);
    const V2 rB = mul(qB, laB - B.lcB);
    V2 uA = (cA + rA) - gA, uB = (cB + rB) - gB;
    float la, lb;
    pulley_axis(&uA, &la);
    pulley_axis(&uB, &lb);
    const float ruA = cross(rA, uA), ruB = cross(rB, uB);
    const float pmA = mA + (iA * ruA) * ruA, pmB = mB + (iB * ruB) * ruB;
    float mass = pmA + (ratio * ratio) * pmB;
    if (mass > 0.0f) mass = 1.0f / mass;
    const float C = (constant - la) - ratio * lb;
    const float impulse = -mass * C;
    const V2 P1 = (-impulse) * uA;
    const V2 P2 = ((-ratio) * impulse) * uB;
    cA = cA + mA * P1;
    aA += iA * cross(rA, P1);
    cB = cB + mB * P2;
    aB += iB * cross(rB, P2);
    okay = fabsf(C) < kLinearSlop;
  } else {  // gear
    const GearBodies G = gear_bodies(j, a);
    const float4 d3 = a.jd3[j];
    const int flags = a.jt[j].y;
    const float ref_a = d3.x, ref_b = d3.y, ratio = d3.z, constant = d3.w;
    const float4 PC = a.pos[G.c], PD = a.pos[G.d];
    V2 cC = v2(PC.x, PC.y), cD = v2(PD.x, PD.y);
    float aC = PC.z, aD = PD.z;
    const Rot qA = rot(aA), qB = rot(aB), qC = rot(aC), qD = rot(aD);
    float coordinate_a, coordinate_b, jw_a, jw_b, jw_c, jw_d, mass = 0.0f;
    V2 jv_ac, jv_bd;
    if (!(flags & 8)) {
      jv_ac = v2(0.0f, 0.0f);
      jw_a = 1.0f;
      jw_c = 1.0f;
      mass += iA + G.iC;
      coordinate_a = (aA - aC) - ref_a;
    } else {
      const V2 axis_c = v2(d2.x, d2.y), laC = v2(d1.x, d1.y);
      const V2 u = mul(qC, axis_c);
      const V2 rC = mul(qC, laC - G.lcC);
      const V2 rA = mul(qA, laA - B.lcA);
      jv_ac = u;
      jw_c = cross(rC, u);
      jw_a = cross(rA, u);
      mass += ((G.mC + mA) + (G.iC * jw_c) * jw_c) + (iA * jw_a) * jw_a;
      const V2 pC = laC - G.lcC;
      const V2 pA = mulT(qC, rA + (cA - cC));
      coordinate_a = dot(pA - pC, axis_c);
    }
    if (!(flags & 16)) {
      jv_bd = v2(0.0f, 0.0f);
      jw_b = ratio;
      jw_d = ratio;
      mass += (ratio * ratio) * (iB + G.iD);
      coordinate_b = (aB - aD) - ref_b;
    } else {
      const V2 axis_d = v2(d2.z, d2.w), laD = v2(d1.z, d1.w);
      const V2 u = mul(qD, axis_d);
      const V2 rD = mul(qD, laD - G.lcD);
      const V2 rB = mul(qB, laB - B.lcB);
      jv_bd = ratio * u;
      jw_d = ratio * cross(rD, u);
      jw_b = ratio * cross(rB, u);
      mass += (((ratio * ratio) * (G.mD + mB)) + (G.iD * jw_d) * jw_d) + (iB * jw_b) * jw_b;
      const V2 pD = laD - G.lcD;
      const V2 pB = mulT(qD, rB + (cB - cD));
      coordinate_b = dot(pB - pD, axis_d);
    }
    const float C = (coordinate_a + ratio * coordinate_b) - constant;
    float impulse = 0.0f;
    if (mass > 0.0f) impulse = -C / mass;
    cA = cA + (mA * impulse) * jv_ac;
    aA += (iA * impulse) * jw_a;
    cB = cB + (mB * impulse) * jv_bd;
    aB += (iB * impulse) * jw_b;
    cC = cC - (G.mC * impulse) * jv_ac;
    aC -= (G.iC * impulse) * jw_c;
    cD = cD - (G.mD * impulse) * jv_bd;
    aD -= (G.iD * impulse) * jw_d;
    if (G.dyn_c) a.pos[G.c] = make_float4(cC.x, cC.y, aC, 0.0f);
    if (G.dyn_d) a.pos[G.d] = make_float4(cD.x, cD.y, aD, 0.0f);
    okay = true;  // the reference's linear error stays 0
  }
  if (B.dyn_a) a.pos[B.a] = make_float4(cA.x, cA.y, aA, 0.0f);
  if (B.dyn_b) a.pos[B.b] = make_float4(cB.x, cB.y, aB, 0.0f);
  return okay;
}

__device__ inline bool extra_kind(int type) {
  return type == B2G_JOINT_FRICTION || type == B2G_JOINT_MOTOR || type == B2G_JOINT_MOUSE ||
         type == B2G_JOINT_WHEEL || type == B2G_JOINT_PULLEY || type == B2G_JOINT_GEAR;
}

// ---- init: B2distanceJoint / B2revoluteJoint::init_velocity_constraints (with warm start) ----
__device__ void joint_init_one(int j, const JointArgs& a) {
  JointBodies B;
  int root;
  if (!joint_bodies(j, a, &B, &root)) return;
  const int type = a.jt[j].x;
  if (extra_kind(type)) {
    extra_init(j, a, B, type);
    return;
  }
  const float4 d0 = a.jd0[j], d1 = a.jd1[j], d2 = a.jd2[j];
  const V2 laA = v2(d0.x, d0.y), laB = v2(d0.z, d0.w);
  const float4 PA = a.pos[B.a], PB = a.pos[B.b];
  const float4 VA = a.vel[B.a], VB = a.vel[B.b];
  V2 vA = v2(VA.x, VA.y), vB = v2(VB.x, VB.y);
  float wA = VA.z, wB = VB.z;
  const Rot qA = rot(PA.z), qB = rot(PB.z);
  const V2 rA = mul(qA, laA - B.lcA);
  const V2 rB = mul(qB, laB - B.lcB);
  float4 imp = a.jimp[j];
  if (type == B2G_JOINT_DISTANCE) {
    const float length = d1.x, min_length = d1.y, max_length = d1.z, stiffness = d1.w, damping = d2.x;
    V2 u = ((v2(PB.x, PB.y) + rB) - v2(PA.x, PA.y)) - rA;
    const float current = len(u);
    float mass = 0.0f;
    if (current > kLinearSlop) {
      u = (1.0f / current) * u;
    } else {
      u = v2(0.0f, 0.0f);
      imp = make_float4(0.0f, 0.0f, 0.0f, imp.w);
    }
    const float crAu = cross(rA, u), crBu = cross(rB, u);
    float inv_mass = ((B.mA + B.iA * crAu * crAu) + B.mB) + B.iB * crBu * crBu;
    mass = inv_mass != 0.0f ? 1.0f / inv_mass : 0.0f;
    float gamma, bias, soft_mass;
    if (stiffness > 0.0f && min_length < max_length) {
      const float C = current - length;
      const float h = a.dt;
      gamma = h * (damping + h * stiffness);
      gamma = gamma != 0.0f ? 1.0f / gamma : 0.0f;
      bias = C * h * stiffness * gamma;
      inv_mass += gamma;
      soft_mass = inv_mass != 0.0f ? 1.0f / inv_mass : 0.0f;
    } else {
      gamma = 0.0f;
      bias = 0.0f;
      soft_mass = mass;
    }
    if (a.warm) {
      imp.x *= a.dt_ratio;
      imp.y *= a.dt_ratio;
      imp.z *= a.dt_ratio;
      const V2 P = ((imp.x + imp.y) - imp.z) * u;
      vA = vA - B.mA * P;
      wA -= B.iA * cross(rA, P);
      vB = vB + B.mB * P;
      wB += B.iB * cross(rB, P);
    } else {
      imp.x = 0.0f;
    }
    a.jk[j] = make_float4(u.x, u.y, mass, soft_mass);
    a.jk2[j] = make_float4(gamma, bias, current, 0.0f);
  } else if (type == B2G_JOINT_PRISMATIC) {
    const int flags = a.jt[j].y;
    const bool limit = flags & 2, motor = flags & 4;
    const float mA = B.mA, mB = B.mB, iA = B.iA, iB = B.iB;
    const V2 cA = v2(PA.x, PA.y), cB = v2(PB.x, PB.y);
    const V2 dd = ((cB - cA) + rB) - rA;
    const V2 axis = mul(qA, v2(d2.y, d2.z));
    const float a1 = cross(dd + rA, axis), a2 = cross(rB, axis);
    float axial = ((mA + mB) + iA * a1 * a1) + iB * a2 * a2;
    if (axial > 0.0f) axial = 1.0f / axial;
    const V2 perp = v2(-axis.y, axis.x);
    const float s1 = cross(dd + rA, perp), s2 = cross(rB, perp);
    const float k11 = ((mA + mB) + iA * s1 * s1) + iB * s2 * s2;
    const float k12 = iA * s1 + iB * s2;
    float k22 = iA + iB;
    if (k22 == 0.0f) k22 = 1.0f;
    float translation = 0.0f;
    float upper = a.jimp2[j];
    if (limit) {
      translation = dot(axis, dd);
    } else {
      imp.w = 0.0f;
      upper = 0.0f;
    }
    if (!motor) imp.z = 0.0f;
    if (a.warm) {
      imp.x *= a.dt_ratio;
      imp.y *= a.dt_ratio;
      imp.z *= a.dt_ratio;
      imp.w *= a.dt_ratio;
      upper *= a.dt_ratio;
      const float axial_impulse = (imp.z + imp.w) - upper;
      const V2 P = imp.x * perp + axial_impulse * axis;
      const float LA = (imp.x * s1 + imp.y) + axial_impulse * a1;
      const float LB = (imp.x * s2 + imp.y) + axial_impulse * a2;
      prismatic_push(B, P, LA, LB, vA, wA, vB, wB);
    } else {
      imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      upper = 0.0f;
    }
    a.jimp2[j] = upper;
    a.jimp[j] = imp;
    a.jr[j] = make_float4(axis.x, axis.y, s1, s2);
    a.jk[j] = make_float4(k11, k12, k22, axial);
    a.jk2[j] = make_float4(a1, a2, translation, 0.0f);
    if (B.dyn_a) a.vel[B.a] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (B.dyn_b) a.vel[B.b] = make_float4(vB.x, vB.y, wB, 0.0f);
    return;
  } else if (type == B2G_JOINT_WELD) {
    const float reference_angle = d1.x, stiffness = d1.y, damping = d1.z;
    const float mA = B.mA, mB = B.mB, iA = B.iA, iB = B.iB;
    const Sym33 K = weld_k(mA, mB, iA, iB, rA, rB);
    Sym33 M;
    float gamma = 0.0f, bias = 0.0f;
    if (stiffness > 0.0f) {
      M = inverse22(K);
      float inv_m = iA + iB;
      const float C = (PB.z - PA.z) - reference_angle;
      const float h = a.dt;
      gamma = h * (damping + h * stiffness);
      gamma = gamma != 0.0f ? 1.0f / gamma : 0.0f;
      bias = C * h * stiffness * gamma;
      inv_m += gamma;
      M.m33 = inv_m != 0.0f ? 1.0f / inv_m : 0.0f;
    } else if (K.m33 == 0.0f) {
      M = inverse22(K);
    } else {
      M = sym_inverse33(K);
    }
    if (a.warm) {
      imp.x *= a.dt_ratio;
      imp.y *= a.dt_ratio;
      imp.z *= a.dt_ratio;
      const V2 P = v2(imp.x, imp.y);
      vA = vA - mA * P;
      wA -= iA * (cross(rA, P) + imp.z);
      vB = vB + mB * P;
      wB += iB * (cross(rB, P) + imp.z);
    } else {
      imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    a.jk[j] = make_float4(M.m11, M.m12, M.m13, M.m22);
    a.jk2[j] = make_float4(M.m23, M.m33, gamma, bias);
  } else {
    const float reference_angle = d1.x;
    const int flags = a.jt[j].y;
    const bool limit = flags & 2, motor = flags & 4;
    const float mA = B.mA, mB = B.mB, iA = B.iA, iB = B.iB;
    const float kxx = ((mA + mB) + rA.y * rA.y * iA) + rB.y * rB.y * iB;
    const float kyx = -rA.y * rA.x * iA - rB.y * rB.x * iB;
    const float kyy = ((mA + mB) + rA.x * rA.x * iA) + rB.x * rB.x * iB;
    float axial = iA + iB;
    bool fixed_rotation;
    if (axial > 0.0f) {
      axial = 1.0f / axial;
      fixed_rotation = false;
    } else {
      fixed_rotation = true;
    }
    const float angle = (PB.z - PA.z) - reference_angle;
    float upper = a.jimp2[j];
    if (!limit || fixed_rotation) {
      imp.w = 0.0f;
      upper = 0.0f;
    }
    if (!motor || fixed_rotation) imp.z = 0.0f;
    if (a.warm) {
      imp.x *= a.dt_ratio;
      imp.y *= a.dt_ratio;
      imp.z *= a.dt_ratio;
      imp.w *= a.dt_ratio;
      upper *= a.dt_ratio;
      const float axial_impulse = (imp.z + imp.w) - upper;
      const V2 P = v2(imp.x, imp.y);
      vA = vA - mA * P;
      wA -= iA * (cross(rA, P) + axial_impulse);
      vB = vB + mB * P;
      wB += iB * (cross(rB, P) + axial_impulse);
    } else {
      imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      upper = 0.0f;
    }
    a.jimp2[j] = upper;
    a.jk[j] = make_float4(kxx, kyx, kyy, axial);
    a.jk2[j] = make_float4(angle, 0.0f, 0.0f, 0.0f);
  }
  a.jimp[j] = imp;
  a.jr[j] = make_float4(rA.x, rA.y, rB.x, rB.y);
  if (B.dyn_a) a.vel[B.a] = make_float4(vA.x, vA.y, wA, 0.0f);
  if (B.dyn_b) a.vel[B.b] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// ---- velocity: solve_velocity_constraints ------------------------------------------------------
__device__ inline void distance_push(float impulse, V2 u, V2 rA, V2 rB, const JointBodies& B, V2& vA, float& wA,
                                     V2& vB, float& wB) {
  const V2 P = impulse * u;
  vA = vA - B.mA * P;
  wA -= B.iA * cross(rA, P);
  vB = vB + B.mB * P;
  wB += B.iB * cross(rB, P);
}

__device__ void joint_velocity_one(int j, const JointArgs& a) {
  JointBodies B;
  int root;
  if (!joint_bodies(j, a, &B, &root)) return;
  const int type = a.jt[j].x;
  if (extra_kind(type)) {
    extra_velocity(j, a, B, type);
    return;
  }
  const float4 VA = a.vel[B.a], VB = a.vel[B.b];
  V2 vA = v2(VA.x, VA.y), vB = v2(VB.x, VB.y);
  float wA = VA.z, wB = VB.z;
  const float4 r = a.jr[j], k = a.jk[j], k2 = a.jk2[j];
  const V2 rA = v2(r.x, r.y), rB = v2(r.z, r.w);
  float4 imp = a.jimp[j];
  if (type == B2G_JOINT_DISTANCE) {
    const float4 d1 = a.jd1[j];
    const float min_length = d1.y, max_length = d1.z, stiffness = d1.w;
    const V2 u = v2(k.x, k.y);
    const float mass = k.z, soft_mass = k.w, gamma = k2.x, bias = k2.y, current = k2.z;
    if (min_length < max_length) {
      if (stiffness > 0.0f) {
        const V2 vpA = vA + cross(wA, rA), vpB = vB + cross(wB, rB);
        const float Cdot = dot(u, vpB - vpA);
        const float impulse = -soft_mass * ((Cdot + bias) + gamma * imp.x);
        imp.x += impulse;
        distance_push(impulse, u, rA, rB, B, vA, wA, vB, wB);
      }
      {
        const float C = current - min_length;
        const float b = bmax(0.0f, C) * a.inv_dt;
        const V2 vpA = vA + cross(wA, rA), vpB = vB + cross(wB, rB);
        const float Cdot = dot(u, vpB - vpA);
        float impulse = -mass * (Cdot + b);
        const float old = imp.y;
        imp.y = bmax(0.0f, imp.y + impulse);
        impulse = imp.y - old;
        distance_push(impulse, u, rA, rB, B, vA, wA, vB, wB);
      }
      {
        const float C = max_length - current;
        const float b = bmax(0.0f, C) * a.inv_dt;
        const V2 vpA = vA + cross(wA, rA), vpB = vB + cross(wB, rB);
        const float Cdot = dot(u, vpA - vpB);
        float impulse = -mass * (Cdot + b);
        const float old = imp.z;
        imp.z = bmax(0.0f, imp.z + impulse);
        impulse = imp.z - old;
        distance_push(-impulse, u, rA, rB, B, vA, wA, vB, wB);
      }
    } else {
      const V2 vpA = vA + cross(wA, rA), vpB = vB + cross(wB, rB);
      const float Cdot = dot(u, vpB - vpA);
      const float impulse = -mass * Cdot;
      imp.x += impulse;
      distance_push(impulse, u, rA, rB, B, vA, wA, vB, wB);
    }
  } else if (type == B2G_JOINT_PRISMATIC) {
    const int flags = a.jt[j].y;
    const bool limit = flags & 2, motor = flags & 4;
    const float4 d1 = a.jd1[j], d2 = a.jd2[j];
    const float lower = d1.y, upper_t = d1.z, max_force = d1.w, motor_speed = d2.x;
    const V2 axis = v2(r.x, r.y), perp = v2(-r.y, r.x);
    const float s1 = r.z, s2 = r.w, k11 = k.x, k12 = k.y, k22 = k.z, axial = k.w;
    const float a1 = k2.x, a2 = k2.y, translation = k2.z;
    float upper = a.jimp2[j];
    if (motor) {
      const float Cdot = (dot(axis, vB - vA) + a2 * wB) - a1 * wA;
      float impulse = axial * (motor_speed - Cdot);
      const float old = imp.z;
      const float max_impulse = a.dt * max_force;
      imp.z = bclamp(imp.z + impulse, -max_impulse, max_impulse);
      impulse = imp.z - old;
      prismatic_push(B, impulse * axis, impulse * a1, impulse * a2, vA, wA, vB, wB);
    }
    if (limit) {
      {
        const float C = translation - lower;
        const float Cdot = (dot(axis, vB - vA) + a2 * wB) - a1 * wA;
        float impulse = -axial * (Cdot + bmax(C, 0.0f) * a.inv_dt);
        const float old = imp.w;
        imp.w = bmax(imp.w + impulse, 0.0f);
        impulse = imp.w - old;
        prismatic_push(B, impulse * axis, impulse * a1, impulse * a2, vA, wA, vB, wB);
      }
      {
        const float C = upper_t - translation;
        const float Cdot = (dot(axis, vA - vB) + a1 * wA) - a2 * wB;
        float impulse = -axial * (Cdot + bmax(C, 0.0f) * a.inv_dt);
        const float old = upper;
        upper = bmax(upper + impulse, 0.0f);
        impulse = upper - old;
        const V2 P = impulse * axis;
        const float LA = impulse * a1, LB = impulse * a2;
        vA = vA + B.mA * P;
        wA += B.iA * LA;
        vB = vB - B.mB * P;
        wB -= B.iB * LB;
      }
    }
    {
      const V2 Cdot = v2((dot(perp, vB - vA) + s2 * wB) - s1 * wA, wB - wA);
      const V2 df = mat22_solve(k11, k12, k12, k22, -Cdot);
      imp.x += df.x;
      imp.y += df.y;
      const V2 P = df.x * perp;
      const float LA = df.x * s1 + df.y;
      const float LB = df.x * s2 + df.y;
      prismatic_push(B, P, LA, LB, vA, wA, vB, wB);
    }
    a.jimp2[j] = upper;
  } else if (type == B2G_JOINT_WELD) {
    const float stiffness = a.jd1[j].y;
    const float mA = B.mA, mB = B.mB, iA = B.iA, iB = B.iB;
    const Sym33 M{k.x, k.y, k.z, k.w, k2.x, k2.y};
    const float gamma = k2.z, bias = k2.w;
    if (stiffness > 0.0f) {
      const float Cdot2 = wB - wA;
      const float impulse2 = -M.m33 * ((Cdot2 + bias) + gamma * imp.z);
      imp.z += impulse2;
      wA -= iA * impulse2;
      wB += iB * impulse2;
      const V2 Cdot1 = ((vB + cross(wB, rB)) - vA) - cross(wA, rA);
      const V2 impulse1 = -mul22(M, Cdot1);
      imp.x += impulse1.x;
      imp.y += impulse1.y;
      const V2 P = impulse1;
      vA = vA - mA * P;
      wA -= iA * cross(rA, P);
      vB = vB + mB * P;
      wB += iB * cross(rB, P);
    } else {
      const V2 Cdot1 = ((vB + cross(wB, rB)) - vA) - cross(wA, rA);
      const float Cdot2 = wB - wA;
      const F3 m = mul33(M, F3{Cdot1.x, Cdot1.y, Cdot2});
      const F3 impulse{-m.x, -m.y, -m.z};
      imp.x += impulse.x;
      imp.y += impulse.y;
      imp.z += impulse.z;
      const V2 P = v2(impulse.x, impulse.y);
      vA = vA - mA * P;
      wA -= iA * (cross(rA, P) + impulse.z);
      vB = vB + mB * P;
      wB += iB * (cross(rB, P) + impulse.z);
    }
  } else {
    const int flags = a.jt[j].y;
    const bool limit = flags & 2, motor = flags & 4;
    const float4 d1 = a.jd1[j], d2 = a.jd2[j];
    const float lower_angle = d1.y, upper_angle = d1.z, max_torque = d1.w, motor_speed = d2.x;
    const float iA = B.iA, iB = B.iB, mA = B.mA, mB = B.mB;
    const float axial = k.w, angle = k2.x;
    const bool fixed_rotation = iA + iB == 0.0f;
    float upper = a.jimp2[j];
    if (motor && !fixed_rotation) {
      const float Cdot = (wB - wA) - motor_speed;
      float impulse = -axial * Cdot;
      const float old = imp.z;
      const float max_impulse = a.dt * max_torque;
      imp.z = bclamp(imp.z + impulse, -max_impulse, max_impulse);
      impulse = imp.z - old;
      wA -= iA * impulse;
      wB += iB * impulse;
    }
    if (limit && !fixed_rotation) {
      {
        const float C = angle - lower_angle;
        const float Cdot = wB - wA;
        float impulse = -axial * (Cdot + bmax(C, 0.0f) * a.inv_dt);
        const float old = imp.w;
        imp.w = bmax(imp.w + impulse, 0.0f);
        impulse = imp.w - old;
        wA -= iA * impulse;
        wB += iB * impulse;
      }
      {
        const float C = upper_angle - angle;
        const float Cdot = wA - wB;
        float impulse = -axial * (Cdot + bmax(C, 0.0f) * a.inv_dt);
        const float old = upper;
        upper = bmax(upper + impulse, 0.0f);
        impulse = upper - old;
        wA += iA * impulse;
        wB -= iB * impulse;
      }
    }
    {
      const V2 Cdot = ((vB + cross(wB, rB)) - vA) - cross(wA, rA);
      const V2 impulse = mat22_solve(k.x, k.y, k.y, k.z, -Cdot);
      imp.x += impulse.x;
      imp.y += impulse.y;
      vA = vA - mA * impulse;
      wA -= iA * cross(rA, impulse);
      vB = vB + mB * impulse;
      wB += iB * cross(rB, impulse);
    }
    a.jimp2[j] = upper;
  }
  a.jimp[j] = imp;
  if (B.dyn_a) a.vel[B.a] = make_float4(vA.x, vA.y, wA, 0.0f);
  if (B.dyn_b) a.vel[B.b] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// ---- position: solve_position_constraints; a joint that is not within tolerance marks its
// island (the island stops iterating only when its contacts and joints are all within it)
__device__ void joint_position_one(int j, const JointArgs& a) {
  JointBodies B;
  int root;
  if (!joint_bodies(j, a, &B, &root)) return;
  if (a.done[root]) return;
  const int type = a.jt[j].x;
  if (extra_kind(type)) {
    if (!extra_position(j, a, B, type)) a.violated[root] = 1;
    return;
  }
  const float4 d0 = a.jd0[j], d1 = a.jd1[j];
  const V2 laA = v2(d0.x, d0.y), laB = v2(d0.z, d0.w);
  const float4 PA = a.pos[B.a], PB = a.pos[B.b];
  V2 cA = v2(PA.x, PA.y), cB = v2(PB.x, PB.y);
  float aA = PA.z, aB = PB.z;
  bool okay;
  if (type == B2G_JOINT_DISTANCE) {
    const float min_length = d1.y, max_length = d1.z;
    const float mass = a.jk[j].z;
    const Rot qA = rot(aA), qB = rot(aB);
    const V2 rA = mul(qA, laA - B.lcA);
    const V2 rB = mul(qB, laB - B.lcB);
    V2 u = ((cB + rB) - cA) - rA;
    const float length = normalize(&u);
    float C;
    bool apply = true;
    if (min_length == max_length) C = length - min_length;
    else if (length < min_length) C = length - min_length;
    else if (max_length < length) C = length - max_length;
    else { C = 0.0f; apply = false; }
    if (apply) {
      const float impulse = -mass * C;
      const V2 P = impulse * u;
      cA = cA - B.mA * P;
      aA -= B.iA * cross(rA, P);
      cB = cB + B.mB * P;
      aB += B.iB * cross(rB, P);
    }
    okay = !apply || fabsf(C) < kLinearSlop;
    if (!apply) return;
  } else if (type == B2G_JOINT_PRISMATIC) {
    const int flags = a.jt[j].y;
    const bool limit = flags & 2;
    const float reference_angle = d1.x, lower = d1.y, upper = d1.z;
    const float4 d2 = a.jd2[j];
    const float mA = B.mA, mB = B.mB, iA = B.iA, iB = B.iB;
    const Rot qA = rot(aA), qB = rot(aB);
    const V2 rA = mul(qA, laA - B.lcA);
    const V2 rB = mul(qB, laB - B.lcB);
    const V2 dd = ((cB + rB) - cA) - rA;
    const V2 axis = mul(qA, v2(d2.y, d2.z));
    const float a1 = cross(dd + rA, axis), a2 = cross(rB, axis);
    const V2 perp = v2(-axis.y, axis.x);
    const float s1 = cross(dd + rA, perp), s2 = cross(rB, perp);
    const V2 C1 = v2(dot(perp, dd), (aB - aA) - reference_angle);
    float linear_error = fabsf(C1.x);
    const float angular_error = fabsf(C1.y);
    bool active = false;
    float C2 = 0.0f;
    if (limit) {
      const float translation = dot(axis, dd);
      if (fabsf(upper - lower) < 2.0f * kLinearSlop) {
        C2 = translation;
        linear_error = bmax(linear_error, fabsf(translation));
        active = true;
      } else if (translation <= lower) {
        C2 = bmin(translation - lower, 0.0f);
        linear_error = bmax(linear_error, lower - translation);
        active = true;
      } else if (translation >= upper) {
        C2 = bmax(translation - upper, 0.0f);
        linear_error = bmax(linear_error, translation - upper);
        active = true;
      }
    }
    F3 impulse;
    const float k11 = ((mA + mB) + iA * s1 * s1) + iB * s2 * s2;
    const float k12 = iA * s1 + iB * s2;
    float k22 = iA + iB;
    if (k22 == 0.0f) k22 = 1.0f;
    if (active) {
      const float k13 = iA * s1 * a1 + iB * s2 * a2;
      const float k23 = iA * a1 + iB * a2;
      const float k33 = ((mA + mB) + iA * a1 * a1) + iB * a2 * a2;
      impulse = solve33(Sym33{k11, k12, k13, k22, k23, k33}, F3{-C1.x, -C1.y, -C2});
    } else {
      const V2 i1 = mat22_solve(k11, k12, k12, k22, -C1);
      impulse = F3{i1.x, i1.y, 0.0f};
    }
    const V2 P = impulse.x * perp + impulse.z * axis;
    const float LA = (impulse.x * s1 + impulse.y) + impulse.z * a1;
    const float LB = (impulse.x * s2 + impulse.y) + impulse.z * a2;
    cA = cA - mA * P;
    aA -= iA * LA;
    cB = cB + mB * P;
    aB += iB * LB;
    okay = linear_error <= kLinearSlop && angular_error <= kAngularSlop;
  } else if (type == B2G_JOINT_WELD) {
    const float reference_angle = d1.x, stiffness = d1.y;
    const float mA = B.mA, mB = B.mB, iA = B.iA, iB = B.iB;
    const Rot qA = rot(aA), qB = rot(aB);
    const V2 rA = mul(qA, laA - B.lcA);
    const V2 rB = mul(qB, laB - B.lcB);
    const Sym33 K = weld_k(mA, mB, iA, iB, rA, rB);
    float position_error, angular_error;
    if (stiffness > 0.0f) {
      const V2 C1 = ((cB + rB) - cA) - rA;
      position_error = len(C1);
      angular_error = 0.0f;
      const V2 P = -mat22_solve(K.m11, K.m12, K.m12, K.m22, C1);
      cA = cA - mA * P;
      aA -= iA * cross(rA, P);
      cB = cB + mB * P;
      aB += iB * cross(rB, P);
    } else {
      const V2 C1 = ((cB + rB) - cA) - rA;
      const float C2 = (aB - aA) - reference_angle;
      position_error = len(C1);
      angular_error = fabsf(C2);
      F3 impulse;
      if (K.m33 > 0.0f) {
        const F3 sol = solve33(K, F3{C1.x, C1.y, C2});
        impulse = F3{-sol.x, -sol.y, -sol.z};
      } else {
        const V2 i2 = -mat22_solve(K.m11, K.m12, K.m12, K.m22, C1);
        impulse = F3{i2.x, i2.y, 0.0f};
      }
      const V2 P = v2(impulse.x, impulse.y);
      cA = cA - mA * P;
      aA -= iA * (cross(rA, P) + impulse.z);
      cB = cB + mB * P;
      aB += iB * (cross(rB, P) + impulse.z);
    }
    okay = position_error <= kLinearSlop && angular_error <= kAngularSlop;
  } else {
    const int flags = a.jt[j].y;
    const bool limit = flags & 2;
    const float reference_angle = d1.x, lower_angle = d1.y, upper_angle = d1.z;
    const float mA = B.mA, mB = B.mB, iA = B.iA, iB = B.iB;
    const float axial = a.jk[j].w;
    float angular_error = 0.0f;
    const bool fixed_rotation = iA + iB == 0.0f;
    if (limit && !fixed_rotation) {
      const float angle = (aB - aA) - reference_angle;
      float C = 0.0f;
      if (fabsf(upper_angle - lower_angle) < 2.0f * kAngularSlop)
        C = bclamp(angle - lower_angle, -kMaxAngularCorrection, kMaxAngularCorrection);
      else if (angle <= lower_angle)
        C = bclamp((angle - lower_angle) + kAngularSlop, -kMaxAngularCorrection, 0.0f);
      else if (angle >= upper_angle)
        C = bclamp((angle - upper_angle) - kAngularSlop, 0.0f, kMaxAngularCorrection);
      const float limit_impulse = -axial * C;
      aA -= iA * limit_impulse;
      aB += iB * limit_impulse;
      angular_error = fabsf(C);
    }
    const Rot qA = rot(aA), qB = rot(aB);
    const V2 rA = mul(qA, laA - B.lcA);
    const V2 rB = mul(qB, laB - B.lcB);
    const V2 C = ((cB + rB) - cA) - rA;
    const float position_error = len(C);
    const float kxx = ((mA + mB) + iA * rA.y * rA.y) + iB * rB.y * rB.y;
    const float kxy = -iA * rA.x * rA.y - iB * rB.x * rB.y;
    const float kyy = ((mA + mB) + iA * rA.x * rA.x) + iB * rB.x * rB.x;
    const V2 impulse = -mat22_solve(kxx, kxy, kxy, kyy, C);
    cA = cA - mA * impulse;
    aA -= iA * cross(rA, impulse);
    cB = cB + mB * impulse;
    aB += iB * cross(rB, impulse);
    okay = position_error <= kLinearSlop && angular_error <= kAngularSlop;
  }
  if (B.dyn_a) a.pos[B.a] = make_float4(cA.x, cA.y, aA, 0.0f);
  if (B.dyn_b) a.pos[B.b] = make_float4(cB.x, cB.y, aB, 0.0f);
  if (!okay) a.violated[root] = 1;
}

__global__ void __launch_bounds__(kBlock) k_joint_init(int begin, int end, JointArgs a) {
  const int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < end) joint_init_one(a.order[i], a);
}
__global__ void __launch_bounds__(kBlock) k_joint_velocity(int begin, int end, JointArgs a) {
  const int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < end) joint_velocity_one(a.order[i], a);
}
__global__ void __launch_bounds__(kBlock) k_joint_position(int begin, int end, JointArgs a) {
  const int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < end) joint_position_one(a.order[i], a);
}
static int blocks(int n) { return (n + kBlock - 1) / kBlock; }

void launch_joint_colours(World* w, int phase, const JointArgs& a) {
  const std::vector<int>& off = w->h_joint_colour_off;
  for (size_t c = 0; c + 1 < off.size(); ++c) {
    const int b0 = off[c], b1 = off[c + 1];
    if (b1 <= b0) continue;
    if (phase == 0)
      hipLaunchKernelGGL(k_joint_init, dim3(blocks(b1 - b0)), dim3(kBlock), 0, w->stream, b0, b1, a);
    else if (phase == 1)
      hipLaunchKernelGGL(k_joint_velocity, dim3(blocks(b1 - b0)), dim3(kBlock), 0, w->stream, b0, b1, a);
    else
      hipLaunchKernelGGL(k_joint_position, dim3(blocks(b1 - b0)), dim3(kBlock), 0, w->stream, b0, b1, a);
  }
}

}  // namespace b2g
