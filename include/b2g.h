/* b2g -- MI355X-native B2world::step for box2d-rs worlds: the C ABI.
 *
 * This is the drop-in boundary. A Rust FFI crate (or any C caller) builds a world with the
 * same definitions box2d-rs uses (B2bodyDef, B2fixtureDef, shapes) and calls b2g_world_step
 * in place of B2world::step. Everything crosses as plain C structs, pointers and sizes.
 *
 * Conventions
 *   - Every function returns an int status: B2G_OK (0) or one of the B2G_ERR_* codes; the
 *     message of the last error of the calling thread is returned by b2g_last_error().
 *   - Caller-owned host buffers with explicit lengths; the library owns all device memory.
 *   - One HIP stream per world; a world must be used by one host thread at a time.
 *   - Citations are to the box2d-rs 0.0.4 sources.
 */
#ifndef B2G_H
#define B2G_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define B2G_OK 0
#define B2G_ERR_INVALID_ARGUMENT 1
#define B2G_ERR_CAPACITY 2
#define B2G_ERR_DEVICE 3
#define B2G_ERR_DEGENERATE 4
#define B2G_ERR_UNSUPPORTED 5

/* Body types: B2bodyType (src/b2_body.rs:20-30). */
#define B2G_STATIC 0
#define B2G_KINEMATIC 1
#define B2G_DYNAMIC 2

/* Shape kinds of b2g_shape_def::type. */
#define B2G_SHAPE_CIRCLE 0  /* params[0..2) = local centre m_p, radius = m_radius        */
#define B2G_SHAPE_EDGE 1    /* params = v1, v2 (two-sided) [+ v0, v3 when flags bit0]     */
#define B2G_SHAPE_POLYGON 2 /* count raw points in params, passed to the hull (set())     */
#define B2G_SHAPE_BOX 4     /* params = hx, hy [, cx, cy, angle when flags bit0]          */
/* A polygon as B2polygonShape stores it (the serde record of src/shapes/serialize_b2_polygon_shape.rs:
 * m_vertices, m_normals, m_centroid), taken verbatim: count vertices in params[0..2*count),
 * normals in params[16..16+2*count), centroid in params[32..34), radius = m_radius. */
#define B2G_SHAPE_POLYGON_EXACT 5

/* Body definition: B2bodyDef (src/b2_body.rs:32-100). 56 bytes. */
typedef struct b2g_body_def {
  int32_t type;
  int32_t allow_sleep, awake, fixed_rotation;
  float position[2];
  float angle;
  float linear_velocity[2];
  float angular_velocity;
  float linear_damping, angular_damping, gravity_scale;
  int32_t bullet; /* B2bodyDef::bullet: continuous collision also against other dynamic bodies */
} b2g_body_def;

/* Shape definition: B2circleShape / B2edgeShape::set_two_sided, set_one_sided /
 * B2polygonShape::set, set_as_box, set_as_box_angle (src/private/collision/b2_*_shape.rs). */
typedef struct b2g_shape_def {
  int32_t type;
  int32_t count;
  int32_t flags;
  float radius;
  float params[36];
} b2g_shape_def;

/* Fixture definition: B2fixtureDef + B2filter (src/b2_fixture.rs:14-58). 32 bytes. A sensor's contacts
 * get no manifold; their touching flag is the shapes' overlap (src/private/dynamics/b2_contact.rs:148-162). */
typedef struct b2g_fixture_def {
  float friction, restitution, restitution_threshold, density;
  int32_t is_sensor;
  uint16_t category_bits, mask_bits;
  int16_t group_index;
  int16_t pad0;
  int32_t pad1;
} b2g_fixture_def;

/* Joint kinds of b2g_joint_def::type: B2jointType values (src/b2_joint.rs:46-58). */
#define B2G_JOINT_DISTANCE 1
#define B2G_JOINT_FRICTION 2
#define B2G_JOINT_GEAR 3
#define B2G_JOINT_MOTOR 4
#define B2G_JOINT_MOUSE 5
#define B2G_JOINT_PRISMATIC 6
#define B2G_JOINT_PULLEY 7
#define B2G_JOINT_REVOLUTE 8
#define B2G_JOINT_WELD 9
#define B2G_JOINT_WHEEL 10

/* Joint definition: B2jointDef (src/b2_joint.rs:149-164) with the fields of its kind; the bodies
 * are creation indices. 156 bytes.
 *   distance  B2distanceJointDef (src/joints/b2_distance_joint.rs): anchors, length, min/max
 *             length, stiffness, damping
 *   revolute  B2revoluteJointDef (src/joints/b2_revolute_joint.rs:41-73): anchors, reference_angle,
 *             limit (lower/upper_angle), motor (motor_speed, max_motor_torque)
 *   weld      B2weldJointDef (src/joints/b2_weld_joint.rs): anchors, reference_angle, stiffness, damping
 *   prismatic B2prismaticJointDef (src/joints/b2_prismatic_joint.rs): anchors, local_axis_a,
 *             reference_angle, lower/upper translation in lower_angle/upper_angle, max motor force
 *             in max_motor_torque, motor_speed
 *   wheel     B2wheelJointDef (src/joints/b2_wheel_joint.rs:39-75): anchors, local_axis_a, lower/upper
 *             translation in lower_angle/upper_angle, max_motor_torque, motor_speed, stiffness, damping
 *   friction  B2frictionJointDef (src/joints/b2_friction_joint.rs): anchors, max_force, max_torque
 *   motor     B2motorJointDef (src/joints/b2_motor_joint.rs): linear_offset, angular_offset,
 *             max_force, max_torque, correction_factor
 *   mouse     B2mouseJointDef (src/joints/b2_mouse_joint.rs:35-52): target, max_force, stiffness,
 *             damping (body B's anchor is the target in body B's frame when the joint is created)
 *   pulley    B2pulleyJointDef (src/joints/b2_pulley_joint.rs): ground anchors, anchors, length_a,
 *             length_b, ratio
 *   gear      B2gearJointDef (src/joints/b2_gear_joint.rs): joint1, joint2 (creation indices of
 *             revolute or prismatic joints; bodies A and B as their bodies B), ratio */
typedef struct b2g_joint_def {
  int32_t type;
  int32_t body_a, body_b;
  int32_t collide_connected;
  float local_anchor_a[2], local_anchor_b[2];
  float length, min_length, max_length, stiffness, damping;                       /* distance (weld: stiffness, damping) */
  float reference_angle, lower_angle, upper_angle, max_motor_torque, motor_speed; /* revolute (weld: reference_angle) */
  int32_t enable_limit, enable_motor;                                             /* revolute, prismatic, wheel */
  float local_axis_a[2];                                                          /* prismatic, wheel */
  float max_force, max_torque;                                                    /* friction, motor, mouse */
  float linear_offset[2], angular_offset, correction_factor;                      /* motor */
  float target[2];                                                                /* mouse */
  float ground_anchor_a[2], ground_anchor_b[2], length_a, length_b, ratio;       /* pulley (gear: ratio) */
  int32_t joint1, joint2;                                                         /* gear */
} b2g_joint_def;

/* Body state as B2body holds it: transform m_xf, sweep, velocities. 64 bytes. */
typedef struct b2g_body_state {
  float p[2], q[2]; /* m_xf.p, (m_xf.q.s, m_xf.q.c) */
  float c[2], a;    /* m_sweep.c, m_sweep.a */
  float v[2], w;    /* m_linear_velocity, m_angular_velocity */
  float c0[2], a0;  /* m_sweep.c0, m_sweep.a0 */
  float sleep_time, awake, pad;
} b2g_body_state;

/* B2manifold (src/b2_collision.rs:51-114) with packed contact-feature ids:
 * id = index_a | index_b << 8 | type_a << 16 | type_b << 24. 72 bytes. */
typedef struct b2g_manifold_point {
  float local_point[2];
  float normal_impulse, tangent_impulse;
  uint32_t id;
  int32_t pad;
} b2g_manifold_point;
typedef struct b2g_manifold {
  int32_t type; /* 0 circles, 1 face A, 2 face B */
  int32_t point_count;
  float local_normal[2], local_point[2];
  b2g_manifold_point points[2];
} b2g_manifold;

/* Per-step profile: B2Profile (src/b2_time_step.rs:5-14) plus device counters. */
typedef struct b2g_profile {
  float step, collide, solve, broadphase; /* milliseconds, device time */
  int32_t pairs, touching, colours, islands, position_iterations;
  int32_t toi_events; /* continuous collision: TOI events solved in the step */
  int32_t island_bins; /* island solve: bins of the last solve (0: per-colour or per-world launches) */
  int32_t pad;
} b2g_profile;

typedef struct b2g_world b2g_world;

/* Message of the last failed call on this thread. */
const char* b2g_last_error(void);
/* Library version string. */
const char* b2g_version(void);

/* B2world::new(gravity) -- src/b2_world.rs:22-30, with its defaults: warm starting, sleeping and
 * continuous physics on (src/private/dynamics/b2_world.rs:37-43); block solve on. */
int b2g_world_create(float gravity_x, float gravity_y, b2g_world** out);
int b2g_world_destroy(b2g_world* world);

/* B2world::set_warm_starting / set_allow_sleeping / set_continuous_physics
 * (src/b2_world.rs:164-185) and G_BLOCK_SOLVE (src/b2_contact.rs:25). continuous runs the TOI
 * pass after each step's solve (src/private/dynamics/b2_world.rs:534-901: times of impact against
 * static and kinematic bodies, and against dynamic bodies for bullets; among equal times of impact
 * the first contact of the GPU's contact-set order goes first). allow_sleep works as in the reference (island sleep
 * after B2_TIME_TO_SLEEP below the sleep tolerances, waking on touch changes; switching it off
 * wakes every body), with one difference: whether a contact is updated in the collide phase
 * depends on its bodies' awake states at the start of the phase, where the reference's
 * sequential pass lets a body woken by an earlier contact of its list update later ones. A
 * world with sleeping bodies cannot record an adjoint tape. */
int b2g_world_set_flags(b2g_world* world, int32_t warm_starting, int32_t allow_sleep, int32_t continuous,
                        int32_t block_solve);

/* B2world::create_body + B2body::create_fixture in bulk (src/private/dynamics/b2_world.rs:83-98,
 * src/private/dynamics/b2_body.rs:155-200). Creates n_bodies bodies, then n_fixtures fixtures in
 * array order, fixture k on body fixture_body[k] (indices relative to the first body created by
 * this call). Writes the broad-phase proxy id of each fixture (the dynamic-tree node index the
 * reference would assign in a create-only world: 0, 1, 3, 5, ...) when proxy_ids is non-null. */
int b2g_world_add(b2g_world* world, int32_t n_bodies, const b2g_body_def* bodies, int32_t n_fixtures,
                  const int32_t* fixture_body, const b2g_shape_def* shapes, const b2g_fixture_def* fixtures,
                  int32_t* proxy_ids);

/* Batched worlds: as b2g_world_add, with a world id per body (0 <= id < 2^20). Bodies of
 * different worlds never interact; every phase of b2g_world_step advances all worlds in the
 * same kernels. No counterpart in the reference (one B2world per world there). */
int b2g_world_add_batched(b2g_world* world, int32_t n_bodies, const b2g_body_def* bodies, const int32_t* body_world,
                          int32_t n_fixtures, const int32_t* fixture_body, const b2g_shape_def* shapes,
                          const b2g_fixture_def* fixtures, int32_t* proxy_ids);
/* Number of worlds (1 + largest world id added). */
int b2g_world_world_count(b2g_world* world, int32_t* out);

/* B2world::step(dt, velocity_iterations, position_iterations) -- src/b2_world.rs:98-100,
 * src/private/dynamics/b2_world.rs:903-959. */
int b2g_world_step(b2g_world* world, float dt, int32_t velocity_iterations, int32_t position_iterations);

/* The three phases of one step, for per-phase parity checks; calling them in order is one
 * b2g_world_step:
 *   B2G_PHASE_COLLIDE : B2contactManager::collide (src/private/dynamics/b2_contact_manager.rs:83-171)
 *   B2G_PHASE_SOLVE   : island solve of all bodies (src/private/dynamics/b2_world.rs:356-507,
 *                       src/private/dynamics/b2_island_private.rs:129-328)
 *   B2G_PHASE_FIND    : synchronize_fixtures + find_new_contacts
 *                       (src/private/dynamics/b2_world.rs:509-528, src/b2_broad_phase.rs:200-249) */
#define B2G_PHASE_COLLIDE 0
#define B2G_PHASE_SOLVE 1
#define B2G_PHASE_FIND 2
int b2g_world_step_phase(b2g_world* world, int32_t phase, float dt, int32_t velocity_iterations,
                         int32_t position_iterations);

int b2g_world_body_count(b2g_world* world, int32_t* out);
int b2g_world_fixture_count(b2g_world* world, int32_t* out);
int b2g_world_contact_count(b2g_world* world, int32_t* out);

/* B2body getters/setters in bulk (get_transform, get_linear_velocity, ...; set_transform
 * and set_linear_velocity semantics without re-synchronising proxies). */
int b2g_world_get_body_states(b2g_world* world, b2g_body_state* out, int32_t n);
int b2g_world_set_body_states(b2g_world* world, const b2g_body_state* in, int32_t n);
/* Mass data per body: mass, inv_mass, I, inv_I, local_center.x, local_center.y. */
int b2g_world_get_mass(b2g_world* world, float* out, int32_t n);
/* B2fixture::set_friction (src/b2_fixture.rs:260-264) for every fixture at once (n = fixture
 * count, frictions >= 0). One difference: a contact's friction is mixed from its fixtures at each
 * step here, so existing contacts take the new values at the next step, where the reference keeps
 * a contact's mixed friction until B2contact::reset_friction. Refused while a tape records. */
int b2g_world_set_fixture_frictions(b2g_world* world, const float* friction, int32_t n);
/* B2body::set_bullet (src/b2_body.rs:504-506): a bullet gets continuous collision also against other
 * dynamic bodies (src/private/dynamics/b2_world.rs:616-621, 770-777). */
int b2g_world_set_bullet(b2g_world* world, int32_t body, int32_t flag);
/* B2world::set_sub_stepping (src/b2_world.rs:187-190): single-stepped continuous physics. A TOI
 * pass ends after its first solved event; until a pass finds no event (step_complete), each step
 * runs its collide and resumes the pass, without a solve (src/private/dynamics/b2_world.rs:545,
 * 676-679, 896-899, 935-939). As in the reference, switching continuous physics off while a pass
 * is unfinished leaves the world without solves. One world only (not batched worlds). */
int b2g_world_set_sub_stepping(b2g_world* world, int32_t flag);
/* B2world::get_sub_stepping (src/b2_world.rs:191-193), and whether the last TOI pass finished. */
int b2g_world_get_sub_stepping(b2g_world* world, int32_t* flag, int32_t* step_complete);

/* Fat AABB of every proxy (B2broadPhase::get_fat_aabb, src/b2_broad_phase.rs) as
 * lower.x, lower.y, upper.x, upper.y. */
int b2g_world_get_fat_aabbs(b2g_world* world, float* out, int32_t n);
/* Overwrites the fat AABB of every proxy (for per-phase parity runs from a given state). */
int b2g_world_set_fat_aabbs(b2g_world* world, const float* in, int32_t n);

/* The contact set (the world contact list, src/b2_contact_manager.rs:67-74), in the device's
 * order: fixture indices A and B (after the registry ordering of
 * src/private/dynamics/b2_contact_registers.rs:67-103), flags (bit0 touching; for a sensor contact
 * the shapes' overlap) and manifolds.
 * *count receives the number of contacts; at most cap are written. */
int b2g_world_get_contacts(b2g_world* world, int32_t cap, int32_t* fixtures_ab, int32_t* flags,
                           b2g_manifold* manifolds, int32_t* count);

/* B2world::create_joint (src/private/dynamics/b2_world.rs:156-275) for n joints, with the
 * constructors' clamping (distance: length, min_length >= linear slop, max_length >= min_length,
 * src/private/dynamics/joints/b2_distance_joint.rs:44-78). Joints join islands, are solved before
 * the contacts in each velocity iteration and after them in each position iteration, and a joint
 * with collide_connected = 0 removes its bodies' contacts from the next step on. A gear joint
 * couples two existing revolute or prismatic joints; its constant, and a mouse joint's anchor,
 * come from the bodies' current poses (src/private/dynamics/joints/b2_gear_joint.rs:8-110,
 * src/joints/b2_mouse_joint.rs:141-150). Not with a tape (b2g_tape_begin refuses joints). */
int b2g_world_add_joints(b2g_world* world, int32_t n, const b2g_joint_def* defs);
/* B2mouseJoint::set_target (src/joints/b2_mouse_joint.rs:114-119): a new target wakes body B. */
int b2g_world_set_joint_target(b2g_world* world, int32_t joint, float x, float y);
/* Warm-start impulses of every joint, 5 floats each (distance: impulse, lower, upper, 0, 0;
 * revolute and prismatic: impulse.x, impulse.y, motor, lower, upper; weld: impulse.x, .y, .z, 0, 0;
 * wheel: impulse, spring, motor, lower, upper; friction and motor: linear.x, linear.y, angular, 0,
 * 0; mouse: impulse.x, impulse.y, 0, 0, 0; pulley and gear: impulse, 0, 0, 0, 0). */
int b2g_world_get_joint_impulses(b2g_world* world, float* out, int32_t n);
/* Overwrites them (same layout): the checkpoint extension of the serde world format restores a
 * world's joints with their impulses. */
int b2g_world_set_joint_impulses(b2g_world* world, const float* in, int32_t n);

/* The shape of every fixture as the device holds it (after the hull, box and edge setters):
 * circles as B2G_SHAPE_CIRCLE, edges as B2G_SHAPE_EDGE (flags bit0 = one-sided), polygons as
 * B2G_SHAPE_POLYGON_EXACT. What the serde world writer records (src/serialize/serialize_b2_fixture.rs:23-52). */
int b2g_world_get_shapes(b2g_world* world, b2g_shape_def* out, int32_t n);

/* Checkpoint of the contact state between steps (the solver-state extension of the serde world
 * format).
 * get_contact_colours: for each contact of b2g_world_get_contacts, in the same order, the solve
 * colour its predecessor had in the last solve (-1: none), which the next step's colouring may
 * keep.
 * get_destroyed_contacts: the touching contacts of the last solve that are not in the current set
 * (their fat AABBs stopped overlapping); the next collide wakes their bodies when sleeping is
 * possible (b2_contact_destroy, src/private/dynamics/b2_contact.rs:33-45). Empty once that collide
 * has run.
 * set_contacts: rebuilds the contact set from the current fat AABBs (the pair search of
 * B2contactManager::find_new_contacts, src/b2_contact_manager.rs:35-45) and gives each contact
 * the manifold (with impulses) and colour of the entry with the same fixtures A, B; contacts
 * without an entry start empty, and entries whose fixtures are not a pair of the rebuilt set are
 * taken as destroyed contacts. Call after b2g_world_set_body_states and b2g_world_set_fat_aabbs.
 * colours may be null (all -1). */
int b2g_world_get_contact_colours(b2g_world* world, int32_t cap, int32_t* colours, int32_t* count);
int b2g_world_get_destroyed_contacts(b2g_world* world, int32_t cap, int32_t* fixtures_ab, b2g_manifold* manifolds,
                                     int32_t* count);
int b2g_world_set_contacts(b2g_world* world, int32_t n, const int32_t* fixtures_ab, const b2g_manifold* manifolds,
                           const int32_t* colours);

/* Step-to-step solver scalars: m_inv_dt0 of B2world (src/private/dynamics/b2_world.rs:924, 949: the
 * warm-start ratio is dt * inv_dt0) and the colour-reuse counters of the coloured solve. */
typedef struct b2g_solver_state {
  float inv_dt0;
  int32_t n_colours, scratch_colours, steps_since_scratch;
} b2g_solver_state;
int b2g_world_get_solver_state(b2g_world* world, b2g_solver_state* out);
int b2g_world_set_solver_state(b2g_world* world, const b2g_solver_state* in);

/* Solver constraint order of the last SOLVE phase: contact indices in solve order (colour by
 * colour) and colour offsets (n_colours + 1 entries). Used to replay the solve on the CPU. */
int b2g_world_get_solver_order(b2g_world* world, int32_t cap, int32_t* contact_order, int32_t* colour_offsets,
                               int32_t colour_cap, int32_t* n_constraints, int32_t* n_colours);

int b2g_world_get_profile(b2g_world* world, b2g_profile* out);

/* Stateless narrowphase: B2contactDynTrait::evaluate (src/b2_contact.rs:97-102) for n pairs of
 * fixtures of this world under the given transforms (4 floats each: p.x, p.y, sin, cos). The
 * pair order is taken as given (A = pairs[2i], B = pairs[2i+1]; A must be the registry's
 * primary shape type). */
int b2g_collide_pairs(b2g_world* world, int32_t n, const int32_t* pairs, const float* xf_a, const float* xf_b,
                      b2g_manifold* out);

/* No counterpart in the reference: b2g_collide_pairs for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, checks and results as b2g_collide_pairs, but its device staging (pairs, both
 * transforms, manifold records) lives in a fixed per-device scratch area of
 * B2G_NOALLOC_MAX_PAIRS pairs (about 7 MB, held by each device the library runs on until the
 * process ends) instead of per-call allocations. n above the maximum: B2G_ERR_CAPACITY. The
 * world must already be on the device (stepped, or queried once through an allocating entry
 * point): otherwise B2G_ERR_INVALID_ARGUMENT. Calls that overlap on one device share the
 * scratch area and race. */
#define B2G_NOALLOC_MAX_PAIRS 65536
int b2g_collide_pairs_noalloc(b2g_world* world, int32_t n, const int32_t* pairs, const float* xf_a,
                              const float* xf_b, b2g_manifold* out);

/* Stateless broadphase: B2broadPhase::update_pairs pair search (src/b2_broad_phase.rs:200-249)
 * over the given fat AABBs of this world's proxies (n = fixture count, 4 floats each), with the
 * world's filters. Writes at most cap pairs (lower, higher fixture index) and the total count. */
int b2g_broadphase_pairs(b2g_world* world, const float* fat_aabbs, int64_t cap, int32_t* pairs, int64_t* count);

/* ---- Adjoint (no counterpart in the reference: box2d-rs has no derivatives) ----
 * Gradients of a run of steps with respect to the body state at its start and the fixture
 * frictions, for callers that optimise through the simulation. The run is recorded by stepping
 * with a tape open; its contact set, manifold types and feature ids, solve order and per-island
 * position-iteration counts are then held fixed (the derivative of the branch the run took), and
 * the backward pass replays segments from body-state checkpoints taken every
 * checkpoint_every steps. While a tape is open, only b2g_world_step may advance the world and
 * no bodies may be added. Worlds with joints, sleeping or continuous physics are refused
 * (B2G_ERR_UNSUPPORTED). */
int b2g_tape_begin(b2g_world* world, int32_t checkpoint_every);
int b2g_tape_end(b2g_world* world);
int b2g_tape_free(b2g_world* world);
int b2g_tape_info(b2g_world* world, int32_t* steps, int64_t* slots);
/* Replays the recorded run with the frozen structure from its start state; writes the final
 * body state, 6 floats per body (c.x, c.y, a, v.x, v.y, w). */
int b2g_tape_replay(b2g_world* world, float* final_state);
/* Reverse pass. cotangent: dL/d(final state), 6 floats per body as above. Writes dL/d(start
 * state) (6 floats per body) and, if grad_friction is not null, dL/d(friction) per fixture. */
int b2g_tape_backward(b2g_world* world, const float* cotangent, float* grad_state, float* grad_friction);
/* One recorded step: per slot 10 ints (body A, body B, fixture A, fixture B, manifold type |
 * point count << 8, carry source of point 0 and 1 as slot * 2 + point over all recorded slots or
 * -1, position iterations run, feature id 0 and 1) and 8 floats (impulses carried in by collide,
 * impulses stored after the solve); colour offsets of the slots; step parameters (dt, dt_ratio,
 * velocity and position iterations, warm starting, block solve). */
int b2g_tape_export(b2g_world* world, int32_t step, int32_t cap, int32_t* slots, float* impulses,
                    int32_t* colour_offsets, int32_t colour_cap, int32_t* n_slots, int32_t* n_colours,
                    float* step_params);

/* Kernels that b2g_world_time_kernel can time, and the item each launch counts. */
enum {
  B2G_KERNEL_NONE = -1,
  B2G_KERNEL_SOLVE_VELOCITY = 0,  /* constraints of one colour, one velocity iteration */
  B2G_KERNEL_SOLVE_POSITION = 1,  /* constraints of one colour, one position iteration */
  B2G_KERNEL_WARM_START = 2,      /* constraints of one colour */
  B2G_KERNEL_INIT_CONSTRAINTS = 3,/* touching contacts */
  B2G_KERNEL_TRAVERSE = 4,        /* proxies in the tree */
  B2G_KERNEL_COLLIDE = 5,         /* contacts (candidate pairs) */
  B2G_KERNEL_COLOR_ROUND = 6,     /* candidate pairs (one colouring round) */
  B2G_KERNEL_COUNT = 7
};

/* Tuning knobs (no counterpart in the reference), set per world by b2g_world_tune:
 *   B2G_TUNE_COLOR_ROUNDS   colouring rounds launched before the first read-back (default 16)
 *   B2G_TUNE_SOLVER_GRAPH   1: the solve's per-colour launches (warm start, velocity and position
 *                           iterations) replay as a captured HIP graph, cached per colour count,
 *                           grid size and buffers; 0: launched one by one (default 0: on MI355X
 *                           the stream of direct launches keeps the GPU busier than the graph's
 *                           replay; timing a kernel with b2g_world_time_kernel also launches them
 *                           one by one)
 *   B2G_TUNE_COLOR_REUSE    1: a contact that persists from the previous step keeps its solve
 *                           colour and only new contacts are coloured; 0: every step colours all
 *                           touching contacts from scratch (default 1). Either way the colouring is
 *                           deterministic; the colour order it gives is the Gauss-Seidel order.
 *   B2G_TUNE_WORLD_SOLVE    batched worlds (b2g_world_add_batched): 1 solves the velocity and the
 *                           position iterations of each world in one workgroup per world, colour
 *                           after colour with workgroup barriers (two launches per step instead of
 *                           one per colour and iteration); 0 launches per colour; -1 (default) is 1
 *                           for batches of at least 16 worlds whose bodies are contiguous per world
 *                           and that have no joints. The constraint order is the colour order
 *                           either way, so the results are the same bit for bit.
 *   B2G_TUNE_WORLD_THREADS_SMALL / _MEDIUM / _LARGE
 *                           threads per world (64, 128, 256 or 512) in that solve for worlds of
 *                           fewer than 64 bodies / 64-255 bodies / 256 or more (defaults 64, 128,
 *                           256; box2d-rs_amd/b2g/autotune.py sweeps them).
 *   B2G_TUNE_ISLAND_SOLVE   1: islands of at most B2G_TUNE_ISLAND_THREADS / 2 constraints are packed
 *                           into bins, and one workgroup per bin runs all velocity (and all position)
 *                           iterations of its islands, colour after colour with workgroup barriers,
 *                           its constraints in registers and its bodies in LDS (two launches per step
 *                           instead of one per colour and iteration); larger islands keep the
 *                           per-colour launches. 0: per-colour launches only. -1 (default): bins
 *                           in steps whose small islands hold at least half of the constraints, unless
 *                           the per-world solve is on. Islands share no dynamic body, so the results
 *                           are the same bit for bit.
 *   B2G_TUNE_ISLAND_THREADS threads per bin of that solve: 64, 128, 256 (default) or 512.
 *   B2G_TUNE_TOI_GLOBAL_ISLAND 1: every TOI island in the global-memory store that islands too large
 *                           for the event workgroup's LDS use (a check of that path; same results);
 *                           0 (default): only those islands. */
#define B2G_TUNE_COLOR_ROUNDS 2
#define B2G_TUNE_SOLVER_GRAPH 3
#define B2G_TUNE_COLOR_REUSE 4
#define B2G_TUNE_WORLD_SOLVE 5
#define B2G_TUNE_WORLD_THREADS_SMALL 6
#define B2G_TUNE_WORLD_THREADS_MEDIUM 7
#define B2G_TUNE_WORLD_THREADS_LARGE 8
#define B2G_TUNE_ISLAND_SOLVE 9
#define B2G_TUNE_ISLAND_THREADS 10
#define B2G_TUNE_TOI_GLOBAL_ISLAND 11
int b2g_world_tune(b2g_world* world, int32_t param, int32_t value);

/* Measurement hook with no counterpart in the reference (its B2Profile times whole phases on
 * the host): records a pair of HIP events on the world's stream around every launch of one
 * kernel (B2G_KERNEL_NONE switches it off). The events add a little time to each timed launch's
 * neighbours, not to its own duration. */
int b2g_world_time_kernel(b2g_world* world, int32_t kernel);
/* Totals of the timed kernel since the last call (which resets them): device milliseconds
 * summed over launches, number of launches, items processed and, if aux is not null, a count
 * known only after the launch (B2G_KERNEL_TRAVERSE: pairs found). Waits for the stream. */
int b2g_world_kernel_time(b2g_world* world, double* ms, int64_t* launches, int64_t* items, int64_t* aux);

#ifdef __cplusplus
}
#endif

#endif /* B2G_H */
