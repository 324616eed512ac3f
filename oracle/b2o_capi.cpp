// C API of the CPU oracle, for the pytest harness (ctypes). TEST INFRASTRUCTURE ONLY: nothing
// in the product library links or loads this file.
//
// Record layouts (all `real` = float in liboracle_f32.so, double in liboracle_f64.so):
//   body def     : see B2oBodyDef        shape def : B2oShapeDef      fixture def : B2oFixtureDef
//   body state   : B2oBodyState (16 reals)
//   manifold     : B2oManifold  (type, count, local normal/point, 2 points with impulses and ids)
#include <algorithm>
#include <chrono>
#include <numeric>
#include <vector>

#include "b2o_world.h"
#include "b2o_toi.h"
#include "b2o_joints.h"

using namespace b2o;

extern "C" {

struct B2oBodyDef {
  int32_t type;
  int32_t allow_sleep, awake, fixed_rotation;
  real position[2];
  real angle;
  real linear_velocity[2];
  real angular_velocity;
  real linear_damping, angular_damping, gravity_scale;
  int32_t bullet;
};

// Shape kinds of B2oShapeDef::type
enum { B2O_CIRCLE = 0, B2O_EDGE = 1, B2O_POLYGON = 2, B2O_BOX = 4, B2O_POLYGON_EXACT = 5 };
struct B2oShapeDef {
  int32_t type;
  int32_t count;   // polygon: number of raw points passed to the hull
  int32_t flags;   // box: bit0 = centred/rotated (set_as_box_angle); edge: bit0 = one-sided
  real radius;     // circle radius
  real params[36]; // circle: p; polygon: points; box: hx,hy,cx,cy,angle; edge: v1,v2,v0,v3;
                   // exact polygon: vertices [0,16), normals [16,32), centroid [32,34)
};

// Joint definition: B2distanceJointDef / B2revoluteJointDef (src/joints/b2_distance_joint.rs,
// src/joints/b2_revolute_joint.rs) with the bodies as creation indices
struct B2oJointDef {
  int32_t type;  // B2jointType: 1 distance, 2 friction, 3 gear, 4 motor, 5 mouse, 6 prismatic, 7 pulley,
                 // 8 revolute, 9 weld, 10 wheel
  int32_t body_a, body_b;
  int32_t collide_connected;
  real local_anchor_a[2], local_anchor_b[2];
  real length, min_length, max_length, stiffness, damping;
  real reference_angle, lower_angle, upper_angle, max_motor_torque, motor_speed;
  int32_t enable_limit, enable_motor;
  real local_axis_a[2];  // prismatic, wheel
  real max_force, max_torque;                                   // friction, motor, mouse
  real linear_offset[2], angular_offset, correction_factor;     // motor
  real target[2];                                               // mouse
  real ground_anchor_a[2], ground_anchor_b[2], length_a, length_b, ratio;  // pulley (gear: ratio)
  int32_t joint1, joint2;                                       // gear
};
// Joint state: the warm-start impulses (distance: impulse, lower, upper; revolute: impulse.x,
// impulse.y, motor, lower, upper)
struct B2oJointState {
  real impulse[5];
  real pad[3];
};

struct B2oFixtureDef {
  real friction, restitution, restitution_threshold, density;
  int32_t is_sensor;
  uint16_t category, mask;
  int16_t group;
  int16_t pad0;
  int32_t pad1;
};

struct B2oBodyState {
  real p[2], q[2];  // transform: position, (sin, cos)
  real c[2], a;     // sweep centre and angle
  real v[2], w;
  real c0[2], a0;
  real sleep_time, awake, pad;
};

struct B2oPoint {
  real local_point[2];
  real normal_impulse, tangent_impulse;
  uint32_t id;
  int32_t pad;
};
struct B2oManifold {
  int32_t type, point_count;
  real local_normal[2], local_point[2];
  B2oPoint points[2];
};

// Full shape record (after hull/box construction) for the stateless collide entry point and
// for inspecting shapes: same meaning as Shape in b2o_core.h.
struct B2oShape {
  int32_t type, count, one_sided, pad;
  real radius;
  real p[2];
  real centroid[2];
  real v0[2], v1[2], v2[2], v3[2];
  real vertices[16];
  real normals[16];
};

int b2o_real_size() { return int(sizeof(real)); }

static void shape_from_def(const B2oShapeDef& d, Shape* s, bool* ok) {
  *ok = true;
  switch (d.type) {
    case B2O_CIRCLE: make_circle(s, Vec2(d.params[0], d.params[1]), d.radius); break;
    case B2O_EDGE:
      if (d.flags & 1)
        make_edge_one_sided(s, Vec2(d.params[4], d.params[5]), Vec2(d.params[0], d.params[1]),
                            Vec2(d.params[2], d.params[3]), Vec2(d.params[6], d.params[7]));
      else
        make_edge_two_sided(s, Vec2(d.params[0], d.params[1]), Vec2(d.params[2], d.params[3]));
      break;
    case B2O_BOX:
      if (d.flags & 1) set_as_box_angle(s, d.params[0], d.params[1], Vec2(d.params[2], d.params[3]), d.params[4]);
      else set_as_box(s, d.params[0], d.params[1]);
      break;
    case B2O_POLYGON: {
      Vec2 pts[kMaxPolygonVertices];
      int n = std::min<int>(d.count, kMaxPolygonVertices);
      for (int i = 0; i < n; ++i) pts[i] = Vec2(d.params[2 * i], d.params[2 * i + 1]);
      *ok = polygon_set(s, pts, n);
      break;
    }
    case B2O_POLYGON_EXACT: {
      // a stored B2polygonShape (the serde record, src/shapes/serialize_b2_polygon_shape.rs:40-90)
      const int n = d.count;
      if (n < 3 || n > kMaxPolygonVertices) { *ok = false; break; }
      s->type = kPolygon;
      s->radius = d.radius;
      s->count = n;
      for (int i = 0; i < n; ++i) {
        s->vertices[i] = Vec2(d.params[2 * i], d.params[2 * i + 1]);
        s->normals[i] = Vec2(d.params[16 + 2 * i], d.params[16 + 2 * i + 1]);
      }
      s->centroid = Vec2(d.params[32], d.params[33]);
      break;
    }
    default: *ok = false;
  }
}

static void shape_to_rec(const Shape& s, B2oShape* r) {
  memset(r, 0, sizeof(*r));
  r->type = s.type;
  r->count = s.count;
  r->one_sided = s.one_sided;
  r->radius = s.radius;
  r->p[0] = s.p.x; r->p[1] = s.p.y;
  r->centroid[0] = s.centroid.x; r->centroid[1] = s.centroid.y;
  r->v0[0] = s.v0.x; r->v0[1] = s.v0.y;
  r->v1[0] = s.v1.x; r->v1[1] = s.v1.y;
  r->v2[0] = s.v2.x; r->v2[1] = s.v2.y;
  r->v3[0] = s.v3.x; r->v3[1] = s.v3.y;
  for (int i = 0; i < s.count; ++i) {
    r->vertices[2 * i] = s.vertices[i].x; r->vertices[2 * i + 1] = s.vertices[i].y;
    r->normals[2 * i] = s.normals[i].x; r->normals[2 * i + 1] = s.normals[i].y;
  }
}

static void shape_from_rec(const B2oShape& r, Shape* s) {
  s->type = r.type;
  s->count = r.count;
  s->one_sided = r.one_sided != 0;
  s->radius = r.radius;
  s->p = Vec2(r.p[0], r.p[1]);
  s->centroid = Vec2(r.centroid[0], r.centroid[1]);
  s->v0 = Vec2(r.v0[0], r.v0[1]);
  s->v1 = Vec2(r.v1[0], r.v1[1]);
  s->v2 = Vec2(r.v2[0], r.v2[1]);
  s->v3 = Vec2(r.v3[0], r.v3[1]);
  for (int i = 0; i < kMaxPolygonVertices; ++i) {
    s->vertices[i] = Vec2(r.vertices[2 * i], r.vertices[2 * i + 1]);
    s->normals[i] = Vec2(r.normals[2 * i], r.normals[2 * i + 1]);
  }
}

static void manifold_to_rec(const Manifold& m, B2oManifold* r) {
  memset(r, 0, sizeof(*r));
  r->type = m.type;
  r->point_count = m.point_count;
  r->local_normal[0] = m.local_normal.x; r->local_normal[1] = m.local_normal.y;
  r->local_point[0] = m.local_point.x; r->local_point[1] = m.local_point.y;
  for (int i = 0; i < 2; ++i) {
    r->points[i].local_point[0] = m.points[i].local_point.x;
    r->points[i].local_point[1] = m.points[i].local_point.y;
    r->points[i].normal_impulse = m.points[i].normal_impulse;
    r->points[i].tangent_impulse = m.points[i].tangent_impulse;
    r->points[i].id = m.points[i].id.key();
  }
}

static void manifold_from_rec(const B2oManifold& r, Manifold* m) {
  m->type = r.type;
  m->point_count = r.point_count;
  m->local_normal = Vec2(r.local_normal[0], r.local_normal[1]);
  m->local_point = Vec2(r.local_point[0], r.local_point[1]);
  for (int i = 0; i < 2; ++i) {
    m->points[i].local_point = Vec2(r.points[i].local_point[0], r.points[i].local_point[1]);
    m->points[i].normal_impulse = r.points[i].normal_impulse;
    m->points[i].tangent_impulse = r.points[i].tangent_impulse;
    uint32_t k = r.points[i].id;
    m->points[i].id.index_a = uint8_t(k & 0xff);
    m->points[i].id.index_b = uint8_t((k >> 8) & 0xff);
    m->points[i].id.type_a = uint8_t((k >> 16) & 0xff);
    m->points[i].id.type_b = uint8_t((k >> 24) & 0xff);
  }
}

static void body_to_state(const Body* b, B2oBodyState* s) {
  s->p[0] = b->xf.p.x; s->p[1] = b->xf.p.y;
  s->q[0] = b->xf.q.s; s->q[1] = b->xf.q.c;
  s->c[0] = b->sweep.c.x; s->c[1] = b->sweep.c.y; s->a = b->sweep.a;
  s->v[0] = b->v.x; s->v[1] = b->v.y; s->w = b->w;
  s->c0[0] = b->sweep.c0.x; s->c0[1] = b->sweep.c0.y; s->a0 = b->sweep.a0;
  s->sleep_time = b->sleep_time;
  s->awake = b->awake ? 1 : 0;
  s->pad = 0;
}

// ---- shapes --------------------------------------------------------------------------------
int b2o_make_shape(const B2oShapeDef* def, B2oShape* out) {
  Shape s;
  bool ok;
  shape_from_def(*def, &s, &ok);
  shape_to_rec(s, out);
  return ok ? 0 : 1;
}

// mass data of a shape record: out = (mass, center.x, center.y, I)
void b2o_shape_mass(const B2oShape* rec, real density, real* out) {
  Shape s;
  shape_from_rec(*rec, &s);
  MassData md;
  compute_mass(s, &md, density);
  out[0] = md.mass; out[1] = md.center.x; out[2] = md.center.y; out[3] = md.I;
}

// AABB of a shape record under transform xf = (p.x, p.y, s, c): out = (lower, upper)
void b2o_shape_aabb(const B2oShape* rec, const real* xf, real* out) {
  Shape s;
  shape_from_rec(*rec, &s);
  Transform t;
  t.p = Vec2(xf[0], xf[1]);
  t.q.s = xf[2]; t.q.c = xf[3];
  AABB a;
  compute_aabb(s, &a, t);
  out[0] = a.lower.x; out[1] = a.lower.y; out[2] = a.upper.x; out[3] = a.upper.y;
}

// ---- stateless narrowphase -----------------------------------------------------------------
// n pairs of shape records (already in registry order: A primary) and transforms
// (p.x, p.y, sin, cos). Writes manifolds without impulses.
void b2o_collide(int n, const B2oShape* shapes_a, const real* xf_a, const B2oShape* shapes_b, const real* xf_b,
                 B2oManifold* out) {
  for (int i = 0; i < n; ++i) {
    Shape a, b;
    shape_from_rec(shapes_a[i], &a);
    shape_from_rec(shapes_b[i], &b);
    Transform ta, tb;
    ta.p = Vec2(xf_a[4 * i], xf_a[4 * i + 1]); ta.q.s = xf_a[4 * i + 2]; ta.q.c = xf_a[4 * i + 3];
    tb.p = Vec2(xf_b[4 * i], xf_b[4 * i + 1]); tb.q.s = xf_b[4 * i + 2]; tb.q.c = xf_b[4 * i + 3];
    Manifold m;
    evaluate(&m, a, ta, b, tb);
    if (m.point_count == 0) m = Manifold();
    manifold_to_rec(m, out + i);
  }
}

// GJK distance of n shape pairs at transforms (p.x, p.y, sin, cos), with or without radii.
// out per pair: distance, point_a (2), point_b (2), iterations.
void b2o_distance(int n, const B2oShape* shapes_a, const real* xf_a, const B2oShape* shapes_b, const real* xf_b,
                  int use_radii, real* out) {
  for (int i = 0; i < n; ++i) {
    Shape a, b;
    shape_from_rec(shapes_a[i], &a);
    shape_from_rec(shapes_b[i], &b);
    Transform ta, tb;
    ta.p = Vec2(xf_a[4 * i], xf_a[4 * i + 1]); ta.q.s = xf_a[4 * i + 2]; ta.q.c = xf_a[4 * i + 3];
    tb.p = Vec2(xf_b[4 * i], xf_b[4 * i + 1]); tb.q.s = xf_b[4 * i + 2]; tb.q.c = xf_b[4 * i + 3];
    DistanceProxy pa, pb;
    pa.set(a);
    pb.set(b);
    SimplexCache cache;
    DistanceOutput d;
    distance(&d, &cache, pa, ta, pb, tb, use_radii != 0);
    real* o = out + 6 * i;
    o[0] = d.distance;
    o[1] = d.pa.x; o[2] = d.pa.y; o[3] = d.pb.x; o[4] = d.pb.y;
    o[5] = real(d.iterations);
  }
}

// Time of impact of n shape pairs over sweeps. A sweep is (local_center (2), c0 (2), c (2), a0, a, alpha0).
// out per pair: state (B2toioutputState: 0 unknown, 1 failed, 2 overlapped, 3 touching, 4 separated), t.
void b2o_time_of_impact(int n, const B2oShape* shapes_a, const real* sweep_a, const B2oShape* shapes_b,
                        const real* sweep_b, real t_max, real* out) {
  for (int i = 0; i < n; ++i) {
    Shape a, b;
    shape_from_rec(shapes_a[i], &a);
    shape_from_rec(shapes_b[i], &b);
    Sweep sw[2];
    const real* src[2] = {sweep_a + 9 * i, sweep_b + 9 * i};
    for (int k = 0; k < 2; ++k) {
      const real* q = src[k];
      sw[k].local_center = Vec2(q[0], q[1]);
      sw[k].c0 = Vec2(q[2], q[3]);
      sw[k].c = Vec2(q[4], q[5]);
      sw[k].a0 = q[6];
      sw[k].a = q[7];
      sw[k].alpha0 = q[8];
    }
    DistanceProxy pa, pb;
    pa.set(a);
    pb.set(b);
    ToiOutput o;
    time_of_impact(&o, pa, pb, sw[0], sw[1], t_max);
    out[2 * i] = real(o.state);
    out[2 * i + 1] = o.t;
  }
}

// World manifold of n manifolds: out per manifold = normal(2), points(4), separations(2).
void b2o_world_manifold(int n, const B2oManifold* ms, const real* xf_a, real* ra, const real* xf_b, real* rb, real* out) {
  for (int i = 0; i < n; ++i) {
    Manifold m;
    manifold_from_rec(ms[i], &m);
    Transform ta, tb;
    ta.p = Vec2(xf_a[4 * i], xf_a[4 * i + 1]); ta.q.s = xf_a[4 * i + 2]; ta.q.c = xf_a[4 * i + 3];
    tb.p = Vec2(xf_b[4 * i], xf_b[4 * i + 1]); tb.q.s = xf_b[4 * i + 2]; tb.q.c = xf_b[4 * i + 3];
    WorldManifold wm;
    world_manifold_initialize(&wm, m, ta, ra[i], tb, rb[i]);
    real* o = out + 8 * i;
    o[0] = wm.normal.x; o[1] = wm.normal.y;
    o[2] = wm.points[0].x; o[3] = wm.points[0].y; o[4] = wm.points[1].x; o[5] = wm.points[1].y;
    o[6] = wm.separations[0]; o[7] = wm.separations[1];
  }
}

// sin/cos exactly as the reference computes a rotation (glibc sinf/cosf for f32)
void b2o_rot(int n, const real* angles, real* sc) {
  for (int i = 0; i < n; ++i) {
    Rot q(angles[i]);
    sc[2 * i] = q.s;
    sc[2 * i + 1] = q.c;
  }
}

// ---- world ---------------------------------------------------------------------------------
void* b2o_world_new(real gx, real gy) { return new World(Vec2(gx, gy)); }
void b2o_world_free(void* w) { delete static_cast<World*>(w); }
void b2o_world_set_flags(void* wp, int warm_starting, int allow_sleep, int block_solve) {
  World* w = static_cast<World*>(wp);
  w->warm_starting = warm_starting != 0;
  w->allow_sleep = allow_sleep != 0;
  w->block_solve = block_solve != 0;
}
void b2o_world_set_continuous(void* wp, int continuous) { static_cast<World*>(wp)->continuous = continuous != 0; }
int b2o_world_toi_events(void* wp) { return static_cast<World*>(wp)->toi_events; }
// B2world::set_sub_stepping (src/b2_world.rs:188-190)
void b2o_world_set_sub_stepping(void* wp, int flag) { static_cast<World*>(wp)->sub_stepping = flag != 0; }
int b2o_world_step_complete(void* wp) { return static_cast<World*>(wp)->step_complete ? 1 : 0; }

// Adds n bodies, then nf fixtures (fixture k on body fixture_body[k], in array order).
// Body indices continue from the existing count. Writes each fixture's proxy id.
// Returns 0, or 1 when a polygon point set is degenerate.
int b2o_world_add(void* wp, int n, const B2oBodyDef* bodies, int nf, const int32_t* fixture_body,
                  const B2oShapeDef* shapes, const B2oFixtureDef* fdefs, int32_t* proxy_ids) {
  World* w = static_cast<World*>(wp);
  int base = int(w->bodies.size());
  for (int i = 0; i < n; ++i) {
    const B2oBodyDef& d = bodies[i];
    w->create_body(d.type, Vec2(d.position[0], d.position[1]), d.angle,
                   Vec2(d.linear_velocity[0], d.linear_velocity[1]), d.angular_velocity, d.linear_damping,
                   d.angular_damping, d.gravity_scale, d.allow_sleep != 0, d.awake != 0, d.fixed_rotation != 0);
    w->bodies.back()->bullet = d.bullet != 0;
  }
  int status = 0;
  for (int k = 0; k < nf; ++k) {
    Shape s;
    bool ok;
    shape_from_def(shapes[k], &s, &ok);
    if (!ok) status = 1;
    Filter f;
    f.category = fdefs[k].category;
    f.mask = fdefs[k].mask;
    f.group = fdefs[k].group;
    Body* b = w->bodies[base + fixture_body[k]];
    Fixture* fx = w->create_fixture(b, s, fdefs[k].density, fdefs[k].friction, fdefs[k].restitution,
                                    fdefs[k].restitution_threshold, f, fdefs[k].is_sensor != 0);
    if (proxy_ids) proxy_ids[k] = fx->proxy_id;
  }
  return status;
}

// B2world::create_joint (src/private/dynamics/b2_world.rs:156-275) with b2_distance_joint_new
// (src/private/dynamics/joints/b2_distance_joint.rs:44-78) / B2revoluteJoint::new
// (src/joints/b2_revolute_joint.rs:253-287); returns the number created
int b2o_world_add_joints(void* wp, int n, const B2oJointDef* defs) {
  World* w = static_cast<World*>(wp);
  const int nb = int(w->bodies.size());
  for (int i = 0; i < n; ++i) {
    const B2oJointDef& d = defs[i];
    if (d.body_a < 0 || d.body_a >= nb || d.body_b < 0 || d.body_b >= nb || d.body_a == d.body_b) return i;
    if (d.type < kDistanceJoint || d.type > kWheelJoint) return i;
    Joint j;
    j.type = d.type;
    j.a = w->bodies[d.body_a];
    j.b = w->bodies[d.body_b];
    j.collide_connected = d.collide_connected != 0;
    j.local_anchor_a = Vec2(d.local_anchor_a[0], d.local_anchor_a[1]);
    j.local_anchor_b = Vec2(d.local_anchor_b[0], d.local_anchor_b[1]);
    if (d.type == kDistanceJoint) {
      j.length = bmax(d.length, kLinearSlop);
      j.min_length = bmax(d.min_length, kLinearSlop);
      j.max_length = bmax(d.max_length, j.min_length);
      j.stiffness = d.stiffness;
      j.damping = d.damping;
    } else if (d.type == kPrismaticJoint) {
      // b2_prismatic_joint::new (src/private/dynamics/joints/b2_prismatic_joint.rs:40-70): the axis
      // normalised, the perpendicular cross(1, axis); limits as translations, motor as a force
      j.local_xaxis_a = Vec2(d.local_axis_a[0], d.local_axis_a[1]);
      j.local_xaxis_a.normalize();
      j.local_yaxis_a = cross(real(1.0), j.local_xaxis_a);
      j.reference_angle = d.reference_angle;
      j.lower_translation = d.lower_angle;
      j.upper_translation = d.upper_angle;
      j.max_motor_force = d.max_motor_torque;
      j.motor_speed = d.motor_speed;
      j.enable_limit = d.enable_limit != 0;
      j.enable_motor = d.enable_motor != 0;
    } else if (d.type == kWeldJoint) {
      // B2weldJoint::new (src/joints/b2_weld_joint.rs): reference angle, stiffness, damping
      j.reference_angle = d.reference_angle;
      j.stiffness = d.stiffness;
      j.damping = d.damping;
    } else if (d.type == kFrictionJoint || d.type == kMotorJoint) {
      // B2frictionJoint::new / B2motorJoint::new (src/joints/b2_friction_joint.rs,
      // src/joints/b2_motor_joint.rs)
      j.max_force = d.max_force;
      j.max_torque = d.max_torque;
      j.linear_offset = Vec2(d.linear_offset[0], d.linear_offset[1]);
      j.angular_offset = d.angular_offset;
      j.correction_factor = d.correction_factor;
    } else if (d.type == kMouseJoint) {
      // B2mouseJoint::new (src/joints/b2_mouse_joint.rs:141-150): body B's anchor at the target
      j.target = Vec2(d.target[0], d.target[1]);
      j.local_anchor_b = mulT(j.b->xf, j.target);
      j.max_force = d.max_force;
      j.stiffness = d.stiffness;
      j.damping = d.damping;
    } else if (d.type == kWheelJoint) {
      // B2wheelJoint::new (src/joints/b2_wheel_joint.rs:329-360): the axis as given, the
      // perpendicular cross(1, axis); limits as translations, motor as a torque, a spring
      j.local_xaxis_a = Vec2(d.local_axis_a[0], d.local_axis_a[1]);
      j.local_yaxis_a = cross(real(1.0), j.local_xaxis_a);
      j.lower_translation = d.lower_angle;
      j.upper_translation = d.upper_angle;
      j.max_motor_torque = d.max_motor_torque;
      j.motor_speed = d.motor_speed;
      j.stiffness = d.stiffness;
      j.damping = d.damping;
      j.enable_limit = d.enable_limit != 0;
      j.enable_motor = d.enable_motor != 0;
    } else if (d.type == kPulleyJoint) {
      // b2_pulley_joint::new (src/private/dynamics/joints/b2_pulley_joint.rs:8-47)
      if (d.ratio == 0) return i;
      j.ground_anchor_a = Vec2(d.ground_anchor_a[0], d.ground_anchor_a[1]);
      j.ground_anchor_b = Vec2(d.ground_anchor_b[0], d.ground_anchor_b[1]);
      j.length_a = d.length_a;
      j.length_b = d.length_b;
      j.ratio = d.ratio;
      j.constant = d.length_a + j.ratio * d.length_b;
    } else if (d.type == kGearJoint) {
      // b2_gear_joint::new (src/private/dynamics/joints/b2_gear_joint.rs:8-110): bodies C and D are
      // the coupled joints' bodies A; the constant from the bodies' current poses
      const int nj = int(w->joints.size());
      if (d.joint1 < 0 || d.joint1 >= nj || d.joint2 < 0 || d.joint2 >= nj) return i;
      const Joint* j1 = w->joints[d.joint1];
      const Joint* j2 = w->joints[d.joint2];
      if ((j1->type != kRevoluteJoint && j1->type != kPrismaticJoint) ||
          (j2->type != kRevoluteJoint && j2->type != kPrismaticJoint))
        return i;
      j.type_a = j1->type;
      j.type_b = j2->type;
      j.c = j1->a;
      j.d = j2->a;
      j.ratio = d.ratio;
      real coordinate_a, coordinate_b;
      {
        const Body* bA = j1->b;
        const Body* bC = j1->a;
        j.local_anchor_c = j1->local_anchor_a;
        j.local_anchor_a = j1->local_anchor_b;
        j.reference_angle_a = j1->reference_angle;
        if (j1->type == kRevoluteJoint) {
          j.local_axis_c.set_zero();
          coordinate_a = bA->sweep.a - bC->sweep.a - j.reference_angle_a;
        } else {
          j.local_axis_c = j1->local_xaxis_a;
          const Vec2 pC = j.local_anchor_c;
          const Vec2 pA = mulT(bC->xf.q, mul(bA->xf.q, j.local_anchor_a) + (bA->xf.p - bC->xf.p));
          coordinate_a = dot(pA - pC, j.local_axis_c);
        }
      }
      {
        const Body* bB = j2->b;
        const Body* bD = j2->a;
        j.local_anchor_d = j2->local_anchor_a;
        j.local_anchor_b = j2->local_anchor_b;
        j.reference_angle_b = j2->reference_angle;
        if (j2->type == kRevoluteJoint) {
          j.local_axis_d.set_zero();
          coordinate_b = bB->sweep.a - bD->sweep.a - j.reference_angle_b;
        } else {
          j.local_axis_d = j2->local_xaxis_a;
          const Vec2 pD = j.local_anchor_d;
          const Vec2 pB = mulT(bD->xf.q, mul(bB->xf.q, j.local_anchor_b) + (bB->xf.p - bD->xf.p));
          coordinate_b = dot(pB - pD, j.local_axis_d);
        }
      }
      j.constant = coordinate_a + j.ratio * coordinate_b;
    } else {
      j.reference_angle = d.reference_angle;
      j.lower_angle = d.lower_angle;
      j.upper_angle = d.upper_angle;
      j.max_motor_torque = d.max_motor_torque;
      j.motor_speed = d.motor_speed;
      j.enable_limit = d.enable_limit != 0;
      j.enable_motor = d.enable_motor != 0;
    }
    w->create_joint(j);
  }
  return n;
}

// B2mouseJoint::set_target (src/joints/b2_mouse_joint.rs:114-119)
void b2o_world_set_joint_target(void* wp, int joint, real x, real y) {
  World* w = static_cast<World*>(wp);
  Joint* j = w->joints[joint];
  const Vec2 t(x, y);
  if (!(t == j->target)) {
    j->b->set_awake(true);
    j->target = t;
  }
}

void b2o_world_get_joints(void* wp, B2oJointState* out) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->joints.size(); ++i) {
    const Joint* j = w->joints[i];
    B2oJointState& o = out[i];
    memset(&o, 0, sizeof(o));
    if (j->type == kDistanceJoint) {
      o.impulse[0] = j->impulse;
      o.impulse[1] = j->lower_impulse;
      o.impulse[2] = j->upper_impulse;
    } else if (j->type == kPrismaticJoint || j->type == kRevoluteJoint) {
      o.impulse[0] = j->impulse2.x;
      o.impulse[1] = j->impulse2.y;
      o.impulse[2] = j->motor_impulse;
      o.impulse[3] = j->lower_impulse;
      o.impulse[4] = j->upper_impulse;
    } else if (j->type == kWeldJoint) {
      o.impulse[0] = j->impulse3[0];
      o.impulse[1] = j->impulse3[1];
      o.impulse[2] = j->impulse3[2];
    } else if (j->type == kFrictionJoint || j->type == kMotorJoint) {
      o.impulse[0] = j->linear_impulse.x;
      o.impulse[1] = j->linear_impulse.y;
      o.impulse[2] = j->angular_impulse;
    } else if (j->type == kMouseJoint) {
      o.impulse[0] = j->impulse2.x;
      o.impulse[1] = j->impulse2.y;
    } else if (j->type == kWheelJoint) {
      o.impulse[0] = j->impulse;
      o.impulse[1] = j->spring_impulse;
      o.impulse[2] = j->motor_impulse;
      o.impulse[3] = j->lower_impulse;
      o.impulse[4] = j->upper_impulse;
    } else if (j->type == kPulleyJoint || j->type == kGearJoint) {
      o.impulse[0] = j->impulse;
    } else {
      o.impulse[0] = j->impulse2.x;
      o.impulse[1] = j->impulse2.y;
      o.impulse[2] = j->motor_impulse;
      o.impulse[3] = j->lower_impulse;
      o.impulse[4] = j->upper_impulse;
    }
  }
}

void b2o_world_step(void* wp, real dt, int vi, int pi) { static_cast<World*>(wp)->step(dt, vi, pi); }

// The collide phase alone (find_new_contacts first when fixtures were added): afterwards the
// contact list is exactly the set of filtered pairs whose fat AABBs overlap.
void b2o_world_collide(void* wp) {
  World* w = static_cast<World*>(wp);
  if (w->new_contacts) {
    w->find_new_contacts();
    w->new_contacts = false;
  }
  w->collide();
}

// Runs `steps` steps and returns the wall time in seconds (for the CPU baseline).
double b2o_world_run(void* wp, real dt, int vi, int pi, int steps) {
  World* w = static_cast<World*>(wp);
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < steps; ++i) w->step(dt, vi, pi);
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int b2o_world_body_count(void* wp) { return int(static_cast<World*>(wp)->bodies.size()); }
int b2o_world_fixture_count(void* wp) { return int(static_cast<World*>(wp)->fixtures.size()); }
int b2o_world_contact_count(void* wp) { return static_cast<World*>(wp)->contact_count; }

void b2o_world_get_bodies(void* wp, B2oBodyState* out) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->bodies.size(); ++i) body_to_state(w->bodies[i], out + i);
}

// body mass data: per body (mass, inv_mass, I, inv_I, local_center.x, local_center.y)
void b2o_world_get_mass(void* wp, real* out) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->bodies.size(); ++i) {
    const Body* b = w->bodies[i];
    real* o = out + 6 * i;
    o[0] = b->mass; o[1] = b->inv_mass; o[2] = b->I; o[3] = b->inv_I;
    o[4] = b->sweep.local_center.x; o[5] = b->sweep.local_center.y;
  }
}

// Overwrites the dynamic state of every body (sweep c, a, velocity; transform re-derived the
// way synchronize_transform does).
// B2body::set_awake (src/b2_body.rs:783-801) on one body
void b2o_world_set_awake(void* wp, int body, int flag) {
  World* w = static_cast<World*>(wp);
  if (body >= 0 && body < int(w->bodies.size())) w->bodies[body]->set_awake(flag != 0);
}

// B2body::set_bullet (src/b2_body.rs:504-506, 775-777)
void b2o_world_set_bullet(void* wp, int body, int flag) {
  World* w = static_cast<World*>(wp);
  if (body >= 0 && body < int(w->bodies.size())) w->bodies[body]->bullet = flag != 0;
}

void b2o_world_set_bodies(void* wp, const B2oBodyState* in) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->bodies.size(); ++i) {
    Body* b = w->bodies[i];
    const B2oBodyState& s = in[i];
    b->sweep.c = Vec2(s.c[0], s.c[1]);
    b->sweep.a = s.a;
    b->sweep.c0 = Vec2(s.c0[0], s.c0[1]);
    b->sweep.a0 = s.a0;
    b->v = Vec2(s.v[0], s.v[1]);
    b->w = s.w;
    b->synchronize_transform();
  }
}

// Per fixture: the tree's fat AABB (lower, upper) and proxy id.
void b2o_world_get_proxies(void* wp, real* fat, int32_t* proxy_id) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->fixtures.size(); ++i) {
    const Fixture* f = w->fixtures[i];
    proxy_id[i] = f->proxy_id;
    if (f->proxy_id >= 0) {
      const AABB& a = w->bp.tree.nodes[f->proxy_id].aabb;
      fat[4 * i] = a.lower.x; fat[4 * i + 1] = a.lower.y; fat[4 * i + 2] = a.upper.x; fat[4 * i + 3] = a.upper.y;
    }
  }
}

// Contacts in world-list order (newest first): fixture indices (A, B after registry swap),
// flags (bit0 touching), manifold, mixed material (friction, restitution, threshold).
int b2o_world_get_contacts(void* wp, int cap, int32_t* fixtures_ab, int32_t* flags, B2oManifold* manifolds, real* material) {
  World* w = static_cast<World*>(wp);
  int i = 0;
  for (Contact* c = w->contact_list; c && i < cap; c = c->next, ++i) {
    fixtures_ab[2 * i] = c->fa->index;
    fixtures_ab[2 * i + 1] = c->fb->index;
    flags[i] = c->touching ? 1 : 0;
    manifold_to_rec(c->manifold, manifolds + i);
    material[3 * i] = c->friction;
    material[3 * i + 1] = c->restitution;
    material[3 * i + 2] = c->threshold;
  }
  return i;
}

// Timing of the last step, milliseconds: step, collide, solve, broadphase.
void b2o_world_get_profile(void* wp, double* out) {
  World* w = static_cast<World*>(wp);
  out[0] = w->profile.step; out[1] = w->profile.collide; out[2] = w->profile.solve; out[3] = w->profile.broadphase;
}

// Pair set by brute force: all fixture pairs (i < j) whose fat AABBs overlap. Independent of
// the tree; used to pin the tree-based pair search. Returns the pair count (<= cap written).
long long b2o_brute_pairs(int n, const real* fat, long long cap, int32_t* out) {
  long long k = 0;
  for (int i = 0; i < n; ++i) {
    AABB a;
    a.lower = Vec2(fat[4 * i], fat[4 * i + 1]); a.upper = Vec2(fat[4 * i + 2], fat[4 * i + 3]);
    for (int j = i + 1; j < n; ++j) {
      AABB b;
      b.lower = Vec2(fat[4 * j], fat[4 * j + 1]); b.upper = Vec2(fat[4 * j + 2], fat[4 * j + 3]);
      if (test_overlap(a, b)) {
        if (k < cap) { out[2 * k] = i; out[2 * k + 1] = j; }
        ++k;
      }
    }
  }
  return k;
}

// ---- solver replay in an explicit constraint order ------------------------------------------
// Solves one step of the island solver (integrate velocities, init, warm start, velocity
// iterations, store impulses, integrate positions, position iterations with per-island early
// exit, copy back) over the given bodies and touching contacts, visiting the constraints of
// each island in the order `order` lists them. Islands are the connected components of the
// contact graph through non-static bodies, as the reference DFS builds them.
//   body_type[nb], body_props[nb*10] = (inv_mass, inv_I, mass, lc.x, lc.y, gravity_scale,
//   linear_damping, angular_damping, force.x, force.y) -- torque fixed at 0 --, state in/out.
//   contacts: body pair, (friction, restitution, threshold, tangent_speed, radiusA, radiusB),
//   manifold in/out (impulses stored back).
// step = (dt, dt_ratio, gx, gy); iters = (velocity, position, warm_starting, block_solve).
// Sleep: body_flags (null: every body awake, may sleep) = awake | allow_sleep << 1 per body, with
// the sleep time in state; an island is solved only if one of its bodies is awake (all of them
// are then set awake, src/private/dynamics/b2_world.rs:389-447), and with allow_sleep the island
// sleep rule of solve_island runs (b2_island_private.rs:291-327).
// Returns the number of islands.
int b2o_solver_replay(int nb, const int32_t* body_type, const real* body_props, B2oBodyState* state,
                      int nc, const int32_t* contact_bodies, const real* contact_params, B2oManifold* manifolds,
                      const int32_t* order, const real* step_params, const int32_t* iters, int32_t* island_of_body,
                      const int32_t* body_flags, int32_t allow_sleep) {
  World w(Vec2(step_params[2], step_params[3]));
  w.allow_sleep = allow_sleep != 0;
  w.block_solve = iters[3] != 0;
  std::vector<Body> bodies(nb);
  for (int i = 0; i < nb; ++i) {
    Body& b = bodies[i];
    const real* pr = body_props + 10 * i;
    b.index = i;
    b.type = body_type[i];
    b.inv_mass = pr[0]; b.inv_I = pr[1]; b.mass = pr[2];
    b.sweep.local_center = Vec2(pr[3], pr[4]);
    b.gravity_scale = pr[5]; b.linear_damping = pr[6]; b.angular_damping = pr[7];
    b.force = Vec2(pr[8], pr[9]);
    b.sweep.c = Vec2(state[i].c[0], state[i].c[1]);
    b.sweep.a = state[i].a;
    b.v = Vec2(state[i].v[0], state[i].v[1]);
    b.w = state[i].w;
    // (bodies of islands that are not solved keep these)
    b.sweep.c0 = Vec2(state[i].c0[0], state[i].c0[1]);
    b.sweep.a0 = state[i].a0;
    b.xf.p = Vec2(state[i].p[0], state[i].p[1]);
    b.xf.q.s = state[i].q[0];
    b.xf.q.c = state[i].q[1];
    b.awake = b.type != kStatic && (body_flags ? (body_flags[i] & 1) != 0 : true);
    b.auto_sleep = body_flags ? (body_flags[i] & 2) != 0 : true;
    b.sleep_time = state[i].sleep_time;
  }
  std::vector<Fixture> fx(2 * nc);
  std::vector<Contact> cs(nc);
  for (int k = 0; k < nc; ++k) {
    fx[2 * k].body = &bodies[contact_bodies[2 * k]];
    fx[2 * k].shape.radius = contact_params[6 * k + 4];
    fx[2 * k + 1].body = &bodies[contact_bodies[2 * k + 1]];
    fx[2 * k + 1].shape.radius = contact_params[6 * k + 5];
    cs[k].fa = &fx[2 * k];
    cs[k].fb = &fx[2 * k + 1];
    cs[k].friction = contact_params[6 * k];
    cs[k].restitution = contact_params[6 * k + 1];
    cs[k].threshold = contact_params[6 * k + 2];
    cs[k].tangent_speed = contact_params[6 * k + 3];
    manifold_from_rec(manifolds[k], &cs[k].manifold);
  }
  // union-find over non-static bodies
  std::vector<int> parent(nb);
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int x) { while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; } return x; };
  for (int k = 0; k < nc; ++k) {
    int a = contact_bodies[2 * k], b = contact_bodies[2 * k + 1];
    if (bodies[a].type == kStatic || bodies[b].type == kStatic) continue;
    int ra = find(a), rb = find(b);
    if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
  }
  std::vector<int> root_island(nb, -1);
  std::vector<IslandIO> islands;
  for (int i = 0; i < nb; ++i) {
    if (bodies[i].type == kStatic) { if (island_of_body) island_of_body[i] = -1; continue; }
    int r = find(i);
    if (root_island[r] < 0) { root_island[r] = int(islands.size()); islands.emplace_back(); }
    islands[root_island[r]].bodies.push_back(&bodies[i]);
    if (island_of_body) island_of_body[i] = root_island[r];
  }
  for (int oi = 0; oi < nc; ++oi) {
    int k = order[oi];
    int a = contact_bodies[2 * k];
    int anchor = bodies[a].type != kStatic ? a : contact_bodies[2 * k + 1];
    islands[root_island[find(anchor)]].contacts.push_back(&cs[k]);
  }
  TimeStep st;
  st.dt = step_params[0];
  st.inv_dt = st.dt > 0 ? real(1.0) / st.dt : real(0);
  st.dt_ratio = step_params[1];
  st.velocity_iterations = iters[0];
  st.position_iterations = iters[1];
  st.warm_starting = iters[2] != 0;
  for (IslandIO& isl : islands) {
    // static bodies touched by this island's contacts get island slots too
    std::vector<Body*> statics;
    for (Contact* c : isl.contacts) {
      for (Body* b : {c->fa->body, c->fb->body})
        if (b->type == kStatic && std::find(statics.begin(), statics.end(), b) == statics.end()) statics.push_back(b);
    }
    bool any_awake = false;
    for (Body* b : isl.bodies) any_awake = any_awake || b->awake;
    if (!any_awake) continue;
    for (Body* b : isl.bodies) b->awake = true;
    for (Body* b : statics) isl.bodies.push_back(b);
    for (size_t i = 0; i < isl.bodies.size(); ++i) isl.bodies[i]->island_index = int(i);
    solve_island(&w, isl, st);
  }
  for (int i = 0; i < nb; ++i) {
    const Body& b = bodies[i];
    if (b.type == kStatic) continue;
    B2oBodyState& s = state[i];
    s.c[0] = b.sweep.c.x; s.c[1] = b.sweep.c.y; s.a = b.sweep.a;
    s.v[0] = b.v.x; s.v[1] = b.v.y; s.w = b.w;
    s.c0[0] = b.sweep.c0.x; s.c0[1] = b.sweep.c0.y; s.a0 = b.sweep.a0;
    s.p[0] = b.xf.p.x; s.p[1] = b.xf.p.y; s.q[0] = b.xf.q.s; s.q[1] = b.xf.q.c;
    s.sleep_time = b.sleep_time;
    s.awake = b.awake ? real(1) : real(0);
  }
  for (int k = 0; k < nc; ++k) manifold_to_rec(cs[k].manifold, manifolds + k);
  return int(islands.size());
}

}  // extern "C"
