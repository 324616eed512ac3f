//! Raw declarations of the b2g C ABI (include/b2g.h), one to one: the MI355X step of box2d-rs's
//! `B2world::step` (src/b2_world.rs:98-100) as hand-written HIP kernels in libb2g.so. The safe
//! wrapper that keeps `B2world::step`'s signature is in `world.rs`.
#![allow(non_camel_case_types)]
use std::os::raw::{c_char, c_int};

pub mod world;

// constants of include/b2g.h
pub const B2G_OK: i32 = 0;
pub const B2G_ERR_INVALID_ARGUMENT: i32 = 1;
pub const B2G_ERR_CAPACITY: i32 = 2;
pub const B2G_ERR_DEVICE: i32 = 3;
pub const B2G_ERR_DEGENERATE: i32 = 4;
pub const B2G_ERR_UNSUPPORTED: i32 = 5;
pub const B2G_STATIC: i32 = 0;
pub const B2G_KINEMATIC: i32 = 1;
pub const B2G_DYNAMIC: i32 = 2;
pub const B2G_SHAPE_CIRCLE: i32 = 0;
pub const B2G_SHAPE_EDGE: i32 = 1;
pub const B2G_SHAPE_POLYGON: i32 = 2;
pub const B2G_SHAPE_BOX: i32 = 4;
pub const B2G_SHAPE_POLYGON_EXACT: i32 = 5;
pub const B2G_JOINT_DISTANCE: i32 = 1;
pub const B2G_JOINT_PRISMATIC: i32 = 6;
pub const B2G_JOINT_REVOLUTE: i32 = 8;
pub const B2G_JOINT_WELD: i32 = 9;
pub const B2G_JOINT_FRICTION: i32 = 2;
pub const B2G_JOINT_GEAR: i32 = 3;
pub const B2G_JOINT_MOTOR: i32 = 4;
pub const B2G_JOINT_MOUSE: i32 = 5;
pub const B2G_JOINT_PULLEY: i32 = 7;
pub const B2G_JOINT_WHEEL: i32 = 10;
pub const B2G_PHASE_COLLIDE: i32 = 0;
pub const B2G_PHASE_SOLVE: i32 = 1;
pub const B2G_PHASE_FIND: i32 = 2;
pub const B2G_NOALLOC_MAX_PAIRS: i32 = 65536;
pub const B2G_TUNE_COLOR_ROUNDS: i32 = 2;
pub const B2G_TUNE_SOLVER_GRAPH: i32 = 3;
pub const B2G_TUNE_COLOR_REUSE: i32 = 4;
pub const B2G_TUNE_WORLD_SOLVE: i32 = 5;
pub const B2G_TUNE_WORLD_THREADS_SMALL: i32 = 6;
pub const B2G_TUNE_WORLD_THREADS_MEDIUM: i32 = 7;
pub const B2G_TUNE_WORLD_THREADS_LARGE: i32 = 8;
pub const B2G_TUNE_ISLAND_SOLVE: i32 = 9;
pub const B2G_TUNE_ISLAND_THREADS: i32 = 10;
pub const B2G_TUNE_TOI_GLOBAL_ISLAND: i32 = 11;
pub const B2G_KERNEL_NONE: i32 = -1;
pub const B2G_KERNEL_SOLVE_VELOCITY: i32 = 0;
pub const B2G_KERNEL_SOLVE_POSITION: i32 = 1;
pub const B2G_KERNEL_WARM_START: i32 = 2;
pub const B2G_KERNEL_INIT_CONSTRAINTS: i32 = 3;
pub const B2G_KERNEL_TRAVERSE: i32 = 4;
pub const B2G_KERNEL_COLLIDE: i32 = 5;
pub const B2G_KERNEL_COLOR_ROUND: i32 = 6;
pub const B2G_KERNEL_COUNT: i32 = 7;

#[repr(C)] pub struct b2g_world { _private: [u8; 0] }

#[repr(C)] #[derive(Clone, Copy, Default)]
pub struct b2g_body_def {
    pub type_: i32, pub allow_sleep: i32, pub awake: i32, pub fixed_rotation: i32,
    pub position: [f32; 2], pub angle: f32, pub linear_velocity: [f32; 2], pub angular_velocity: f32,
    pub linear_damping: f32, pub angular_damping: f32, pub gravity_scale: f32, pub bullet: i32,
}
#[repr(C)] #[derive(Clone, Copy)]
pub struct b2g_shape_def { pub type_: i32, pub count: i32, pub flags: i32, pub radius: f32, pub params: [f32; 36] }
#[repr(C)] #[derive(Clone, Copy, Default)]
pub struct b2g_fixture_def {
    pub friction: f32, pub restitution: f32, pub restitution_threshold: f32, pub density: f32,
    pub is_sensor: i32, pub category_bits: u16, pub mask_bits: u16, pub group_index: i16, pub pad0: i16, pub pad1: i32,
}
#[repr(C)] #[derive(Clone, Copy, Default)]
pub struct b2g_body_state {
    pub p: [f32; 2], pub q: [f32; 2], pub c: [f32; 2], pub a: f32, pub v: [f32; 2], pub w: f32,
    pub c0: [f32; 2], pub a0: f32, pub sleep_time: f32, pub awake: f32, pub pad: f32,
}

#[repr(C)] #[derive(Clone, Copy, Default)]
pub struct b2g_joint_def {
    pub type_: i32, pub body_a: i32, pub body_b: i32, pub collide_connected: i32,
    pub local_anchor_a: [f32; 2], pub local_anchor_b: [f32; 2],
    pub length: f32, pub min_length: f32, pub max_length: f32, pub stiffness: f32, pub damping: f32,
    pub reference_angle: f32, pub lower_angle: f32, pub upper_angle: f32, pub max_motor_torque: f32,
    pub motor_speed: f32, pub enable_limit: i32, pub enable_motor: i32, pub local_axis_a: [f32; 2],
    pub max_force: f32, pub max_torque: f32, pub linear_offset: [f32; 2], pub angular_offset: f32,
    pub correction_factor: f32, pub target: [f32; 2], pub ground_anchor_a: [f32; 2], pub ground_anchor_b: [f32; 2],
    pub length_a: f32, pub length_b: f32, pub ratio: f32, pub joint1: i32, pub joint2: i32,
}
#[repr(C)] #[derive(Clone, Copy, Default)]
pub struct b2g_manifold_point { pub local_point: [f32; 2], pub normal_impulse: f32, pub tangent_impulse: f32, pub id: u32, pub pad: i32 }
#[repr(C)] #[derive(Clone, Copy, Default)]
pub struct b2g_manifold {
    pub type_: i32, pub point_count: i32, pub local_normal: [f32; 2], pub local_point: [f32; 2],
    pub points: [b2g_manifold_point; 2],
}
#[repr(C)] #[derive(Clone, Copy, Default)]
pub struct b2g_profile {
    pub step: f32, pub collide: f32, pub solve: f32, pub broadphase: f32,
    pub pairs: i32, pub touching: i32, pub colours: i32, pub islands: i32, pub position_iterations: i32,
    pub toi_events: i32, pub island_bins: i32, pub pad: i32,
}
#[repr(C)] #[derive(Clone, Copy, Default)]
pub struct b2g_solver_state { pub inv_dt0: f32, pub n_colours: i32, pub scratch_colours: i32, pub steps_since_scratch: i32 }

extern "C" {
    pub fn b2g_last_error() -> *const c_char;
    pub fn b2g_version() -> *const c_char;
    pub fn b2g_world_create(gx: f32, gy: f32, out: *mut *mut b2g_world) -> c_int;
    pub fn b2g_world_destroy(w: *mut b2g_world) -> c_int;
    pub fn b2g_world_set_flags(w: *mut b2g_world, warm: i32, sleep: i32, continuous: i32, block: i32) -> c_int;
    pub fn b2g_world_add(w: *mut b2g_world, nb: i32, bodies: *const b2g_body_def, nf: i32,
                         fixture_body: *const i32, shapes: *const b2g_shape_def,
                         fixtures: *const b2g_fixture_def, proxy_ids: *mut i32) -> c_int;
    pub fn b2g_world_add_batched(w: *mut b2g_world, nb: i32, bodies: *const b2g_body_def, body_world: *const i32,
                                 nf: i32, fixture_body: *const i32, shapes: *const b2g_shape_def,
                                 fixtures: *const b2g_fixture_def, proxy_ids: *mut i32) -> c_int;
    pub fn b2g_world_world_count(w: *mut b2g_world, out: *mut i32) -> c_int;
    pub fn b2g_world_step(w: *mut b2g_world, dt: f32, vi: i32, pi: i32) -> c_int;
    pub fn b2g_world_step_phase(w: *mut b2g_world, phase: i32, dt: f32, vi: i32, pi: i32) -> c_int;
    pub fn b2g_world_body_count(w: *mut b2g_world, out: *mut i32) -> c_int;
    pub fn b2g_world_fixture_count(w: *mut b2g_world, out: *mut i32) -> c_int;
    pub fn b2g_world_contact_count(w: *mut b2g_world, out: *mut i32) -> c_int;
    pub fn b2g_world_get_body_states(w: *mut b2g_world, out: *mut b2g_body_state, n: i32) -> c_int;
    pub fn b2g_world_set_body_states(w: *mut b2g_world, inp: *const b2g_body_state, n: i32) -> c_int;
    pub fn b2g_world_get_mass(w: *mut b2g_world, out: *mut f32, n: i32) -> c_int;
    pub fn b2g_world_set_fixture_frictions(w: *mut b2g_world, friction: *const f32, n: i32) -> c_int;
    pub fn b2g_world_set_bullet(w: *mut b2g_world, body: i32, flag: i32) -> c_int;
    pub fn b2g_world_set_sub_stepping(w: *mut b2g_world, flag: i32) -> c_int;
    pub fn b2g_world_get_sub_stepping(w: *mut b2g_world, flag: *mut i32, step_complete: *mut i32) -> c_int;
    pub fn b2g_world_get_fat_aabbs(w: *mut b2g_world, out: *mut f32, n: i32) -> c_int;
    pub fn b2g_world_set_fat_aabbs(w: *mut b2g_world, inp: *const f32, n: i32) -> c_int;
    pub fn b2g_world_get_contacts(w: *mut b2g_world, cap: i32, fixtures_ab: *mut i32, flags: *mut i32,
                                  manifolds: *mut b2g_manifold, count: *mut i32) -> c_int;
    pub fn b2g_world_add_joints(w: *mut b2g_world, n: i32, defs: *const b2g_joint_def) -> c_int;
    pub fn b2g_world_set_joint_target(w: *mut b2g_world, joint: i32, x: f32, y: f32) -> c_int;
    pub fn b2g_world_get_joint_impulses(w: *mut b2g_world, out: *mut f32, n: i32) -> c_int;
    pub fn b2g_world_set_joint_impulses(w: *mut b2g_world, inp: *const f32, n: i32) -> c_int;
    pub fn b2g_world_get_shapes(w: *mut b2g_world, out: *mut b2g_shape_def, n: i32) -> c_int;
    pub fn b2g_world_get_contact_colours(w: *mut b2g_world, cap: i32, colours: *mut i32, count: *mut i32) -> c_int;
    pub fn b2g_world_get_destroyed_contacts(w: *mut b2g_world, cap: i32, fixtures_ab: *mut i32,
                                            manifolds: *mut b2g_manifold, count: *mut i32) -> c_int;
    pub fn b2g_world_set_contacts(w: *mut b2g_world, n: i32, fixtures_ab: *const i32,
                                  manifolds: *const b2g_manifold, colours: *const i32) -> c_int;
    pub fn b2g_world_get_solver_state(w: *mut b2g_world, out: *mut b2g_solver_state) -> c_int;
    pub fn b2g_world_set_solver_state(w: *mut b2g_world, inp: *const b2g_solver_state) -> c_int;
    pub fn b2g_world_get_solver_order(w: *mut b2g_world, cap: i32, contact_order: *mut i32, colour_offsets: *mut i32,
                                      colour_cap: i32, n_constraints: *mut i32, n_colours: *mut i32) -> c_int;
    pub fn b2g_world_get_profile(w: *mut b2g_world, out: *mut b2g_profile) -> c_int;
    pub fn b2g_collide_pairs(w: *mut b2g_world, n: i32, pairs: *const i32, xf_a: *const f32, xf_b: *const f32,
                             out: *mut b2g_manifold) -> c_int;
    pub fn b2g_collide_pairs_noalloc(w: *mut b2g_world, n: i32, pairs: *const i32, xf_a: *const f32,
                                     xf_b: *const f32, out: *mut b2g_manifold) -> c_int;
    pub fn b2g_broadphase_pairs(w: *mut b2g_world, fat_aabbs: *const f32, cap: i64, pairs: *mut i32,
                                count: *mut i64) -> c_int;
    pub fn b2g_tape_begin(w: *mut b2g_world, checkpoint_every: i32) -> c_int;
    pub fn b2g_tape_end(w: *mut b2g_world) -> c_int;
    pub fn b2g_tape_free(w: *mut b2g_world) -> c_int;
    pub fn b2g_tape_info(w: *mut b2g_world, steps: *mut i32, slots: *mut i64) -> c_int;
    pub fn b2g_tape_replay(w: *mut b2g_world, final_state: *mut f32) -> c_int;
    pub fn b2g_tape_backward(w: *mut b2g_world, cotangent: *const f32, grad_state: *mut f32,
                             grad_friction: *mut f32) -> c_int;
    pub fn b2g_tape_export(w: *mut b2g_world, step: i32, cap: i32, slots: *mut i32, impulses: *mut f32,
                           colour_offsets: *mut i32, colour_cap: i32, n_slots: *mut i32, n_colours: *mut i32,
                           step_params: *mut f32) -> c_int;
    pub fn b2g_world_tune(w: *mut b2g_world, param: i32, value: i32) -> c_int;
    pub fn b2g_world_time_kernel(w: *mut b2g_world, kernel: i32) -> c_int;
    pub fn b2g_world_kernel_time(w: *mut b2g_world, ms: *mut f64, launches: *mut i64, items: *mut i64,
                                 aux: *mut i64) -> c_int;
}
