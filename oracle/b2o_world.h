// CPU oracle, part 2: dynamic tree, broadphase, contacts, island solver and world step.
// TEST INFRASTRUCTURE ONLY (see b2o_core.h). Sequential restatement of the reference in its
// own order: body/contact lists iterate newest-first, the tree is restated exactly (so proxy
// ids and pair order match), islands are built by the reference DFS.
#pragma once
#include <vector>
#include "b2o_core.h"

namespace b2o {

static const int kNullNode = -1;

// ---- dynamic tree: src/private/collision/b2_dynamic_tree.rs:7-489 -----------------------
struct TreeNode {
  AABB aabb;
  int user = -1;        // fixture index for leaves
  int parent = kNullNode;  // doubles as "next" in the free list
  int child1 = kNullNode, child2 = kNullNode;
  int height = -1;
  bool moved = false;
  bool is_leaf() const { return child1 == kNullNode; }
};

struct DynamicTree {
  int root = kNullNode;
  std::vector<TreeNode> nodes;
  int node_count = 0, node_capacity = 16, free_list = 0, insertion_count = 0;

  DynamicTree() {
    nodes.resize(node_capacity);
    for (int i = 0; i < node_capacity - 1; ++i) { nodes[i].parent = i + 1; nodes[i].height = -1; }
    nodes[node_capacity - 1].parent = kNullNode;
    nodes[node_capacity - 1].height = -1;
  }
  int allocate_node();
  void free_node(int id);
  int create_proxy(const AABB& aabb, int user);
  bool move_proxy(int id, const AABB& aabb, Vec2 displacement);
  void insert_leaf(int leaf);
  void remove_leaf(int leaf);
  int balance(int iA);
  template <class F> void query(F&& cb, const AABB& aabb) const;  // src/b2_dynamic_tree.rs:239-267
};

template <class F> void DynamicTree::query(F&& cb, const AABB& aabb) const {
  std::vector<int> stack;
  stack.reserve(256);
  stack.push_back(root);
  while (!stack.empty()) {
    int id = stack.back();
    stack.pop_back();
    if (id == kNullNode) continue;
    const TreeNode& n = nodes[id];
    if (test_overlap(n.aabb, aabb)) {
      if (n.is_leaf()) {
        if (!cb(id)) return;
      } else {
        stack.push_back(n.child1);
        stack.push_back(n.child2);
      }
    }
  }
}

// ---- broadphase: src/private/collision/b2_broad_phase.rs, src/b2_broad_phase.rs:200-249 ---
struct Pair { int a, b; };
struct BroadPhase {
  DynamicTree tree;
  std::vector<int> move_buffer;
  std::vector<Pair> pairs;
  int create_proxy(const AABB& aabb, int user) {
    int id = tree.create_proxy(aabb, user);
    move_buffer.push_back(id);
    return id;
  }
  void move_proxy(int id, const AABB& aabb, Vec2 d) {
    if (tree.move_proxy(id, aabb, d)) move_buffer.push_back(id);
  }
  template <class F> void update_pairs(F&& add_pair);
};

template <class F> void BroadPhase::update_pairs(F&& add_pair) {
  pairs.clear();
  for (size_t i = 0; i < move_buffer.size(); ++i) {
    int q = move_buffer[i];
    if (q == kNullNode) continue;
    AABB fat = tree.nodes[q].aabb;
    tree.query([&](int pid) -> bool {
      if (pid == q) return true;
      bool moved = tree.nodes[pid].moved;
      if (moved && pid > q) return true;
      pairs.push_back(Pair{bmin(pid, q), bmax(pid, q)});
      return true;
    }, fat);
  }
  for (const Pair& p : pairs) add_pair(tree.nodes[p.a].user, tree.nodes[p.b].user);
  for (int id : move_buffer) if (id != kNullNode) tree.nodes[id].moved = false;
  move_buffer.clear();
}

// ---- bodies, fixtures, contacts --------------------------------------------------------
enum BodyType { kStatic = 0, kKinematic = 1, kDynamic = 2 };

struct Contact;
struct Body;
// src/b2_contact.rs B2contactEdge; lists are push_front / newest-first
struct ContactEdge {
  Body* other = nullptr;
  Contact* contact = nullptr;
  ContactEdge *prev = nullptr, *next = nullptr;
};

struct Filter { uint16_t category = 1, mask = 0xFFFF; int16_t group = 0; };

// Joints: B2jointType values (src/b2_joint.rs:46-58)
enum JointType {
  kDistanceJoint = 1,
  kFrictionJoint = 2,
  kGearJoint = 3,
  kMotorJoint = 4,
  kMouseJoint = 5,
  kPrismaticJoint = 6,
  kPulleyJoint = 7,
  kRevoluteJoint = 8,
  kWeldJoint = 9,
  kWheelJoint = 10
};
struct Joint;
// src/b2_joint.rs B2jointEdge; a body's list is push_front / newest-first
struct JointEdge {
  Body* other = nullptr;
  Joint* joint = nullptr;
  JointEdge *prev = nullptr, *next = nullptr;
};

struct Fixture {
  int index = -1;
  Body* body = nullptr;
  Shape shape;
  real density = 0, friction = real(0.2), restitution = 0, threshold = real(1.0);
  bool sensor = false;
  Filter filter;
  int proxy_id = -1;
  AABB aabb;
};

struct Body {
  int index = -1;
  int type = kStatic;
  bool island_flag = false, awake = false, auto_sleep = true, fixed_rotation = false, enabled = true;
  bool bullet = false;  // B2bodyDef::bullet (src/private/dynamics/b2_body.rs:25-27)
  Transform xf;
  Sweep sweep;
  Vec2 v;
  real w = 0;
  Vec2 force;
  real torque = 0;
  real mass = 0, inv_mass = 0, I = 0, inv_I = 0;
  real linear_damping = 0, angular_damping = 0, gravity_scale = 1;
  real sleep_time = 0;
  int island_index = -1;
  std::vector<Fixture*> fixtures;  // newest first (src/private/dynamics/b2_body.rs:183)
  ContactEdge* contact_list = nullptr;
  JointEdge* joint_list = nullptr;
  void synchronize_transform() {
    xf.q.set(sweep.a);
    xf.p = sweep.c - mul(xf.q, sweep.local_center);
  }
  void set_awake(bool flag);
  // src/b2_body.rs:979-986: advance to a safe time without touching the broad-phase
  void advance(real alpha);
};

struct Contact {
  bool island_flag = false, touching = false, enabled = true, filter = false;
  // continuous collision: cached time of impact (src/private/dynamics/b2_contact.rs E_TOI_FLAG, m_toi, m_toi_count)
  bool toi_flag = false;
  int toi_count = 0;
  real toi = 1;
  Fixture *fa = nullptr, *fb = nullptr;
  Manifold manifold;
  real friction = 0, restitution = 0, threshold = 0, tangent_speed = 0;
  ContactEdge node_a, node_b;
  Contact *prev = nullptr, *next = nullptr;
};

// src/b2_joint.rs B2joint + the distance / revolute joint state
// (src/joints/b2_distance_joint.rs, src/joints/b2_revolute_joint.rs)
struct Joint {
  int type = kRevoluteJoint;
  int index = -1;
  Body *a = nullptr, *b = nullptr;
  bool collide_connected = false, island_flag = false;
  JointEdge edge_a, edge_b;
  Vec2 local_anchor_a, local_anchor_b;
  // distance
  real length = 1, min_length = 0, max_length = kMaxFloat, stiffness = 0, damping = 0;
  real gamma = 0, bias = 0, impulse = 0, lower_impulse = 0, upper_impulse = 0, current_length = 0, soft_mass = 0;
  Vec2 u;
  // revolute
  real reference_angle = 0, lower_angle = 0, upper_angle = 0, max_motor_torque = 0, motor_speed = 0;
  bool enable_limit = false, enable_motor = false;
  Vec2 impulse2;
  real motor_impulse = 0, axial_mass = 0, angle = 0;
  Mat22 K;
  // prismatic (src/joints/b2_prismatic_joint.rs): axes in A's frame, translation limits, motor force
  Vec2 local_xaxis_a, local_yaxis_a, axis, perp;
  real lower_translation = 0, upper_translation = 0, max_motor_force = 0;
  real s1 = 0, s2 = 0, a1 = 0, a2 = 0, translation = 0;
  // weld (src/joints/b2_weld_joint.rs): impulse (x, y, angular), the 3x3 mass (symmetric)
  real impulse3[3] = {0, 0, 0};
  real mass33[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // [column][row]
  // friction and motor (src/joints/b2_friction_joint.rs, b2_motor_joint.rs): linear and angular
  // impulses and masses, the motor's offsets, errors and correction factor
  real max_force = 0, max_torque = 0, angular_impulse = 0, angular_mass = 0;
  Vec2 linear_impulse, linear_offset, linear_error;
  real angular_offset = 0, angular_error = 0, correction_factor = 0;
  Mat22 linear_mass;
  // mouse (src/joints/b2_mouse_joint.rs): target, soft point constraint (impulse in impulse2)
  Vec2 target, mouse_c;
  real beta = 0;
  Mat22 mouse_mass;
  // wheel (src/joints/b2_wheel_joint.rs): axes in A's frame, spring, motor torque
  Vec2 ax, ay;
  real s_ax = 0, s_bx = 0, s_ay = 0, s_by = 0, spring_mass = 0, spring_impulse = 0, motor_mass = 0;
  // pulley (src/joints/b2_pulley_joint.rs)
  Vec2 ground_anchor_a, ground_anchor_b, u_a, u_b;
  real length_a = 0, length_b = 0, ratio = 1, constant = 0;
  // gear (src/joints/b2_gear_joint.rs): bodies C and D of the coupled joints, their frames
  Body *c = nullptr, *d = nullptr;
  int type_a = 0, type_b = 0, index_c = 0, index_d = 0;
  Vec2 local_anchor_c, local_anchor_d, local_axis_c, local_axis_d, lc_a, lc_b, lc_c, lc_d, jv_ac, jv_bd;
  real reference_angle_a = 0, reference_angle_b = 0;
  real m_a = 0, m_b = 0, m_c = 0, m_d = 0, i_a = 0, i_b = 0, i_c = 0, i_d = 0;
  real jw_a = 0, jw_b = 0, jw_c = 0, jw_d = 0;
  // solver temporaries
  int index_a = 0, index_b = 0;
  Vec2 local_center_a, local_center_b, rA, rB;
  real inv_mass_a = 0, inv_mass_b = 0, inv_ia = 0, inv_ib = 0, mass = 0;
};

// ---- contact solver: src/private/dynamics/b2_contact_solver{,_private}.rs ----------------
struct VelocityConstraintPoint {
  Vec2 rA, rB;
  real normal_impulse = 0, tangent_impulse = 0, normal_mass = 0, tangent_mass = 0, velocity_bias = 0;
};
struct VelocityConstraint {
  VelocityConstraintPoint points[kMaxManifoldPoints];
  Vec2 normal;
  Mat22 normal_mass, K;
  int indexA = 0, indexB = 0;
  real invMassA = 0, invMassB = 0, invIA = 0, invIB = 0;
  real friction = 0, restitution = 0, threshold = 0, tangent_speed = 0;
  int point_count = 0, contact_index = 0;
};
struct PositionConstraint {
  Vec2 local_points[kMaxManifoldPoints];
  Vec2 local_normal, local_point;
  int indexA = 0, indexB = 0;
  real invMassA = 0, invMassB = 0;
  Vec2 localCenterA, localCenterB;
  real invIA = 0, invIB = 0;
  int type = kCircles;
  real radiusA = 0, radiusB = 0;
  int point_count = 0;
};
struct Position { Vec2 c; real a = 0; };
struct Velocity { Vec2 v; real w = 0; };
struct TimeStep { real dt = 0, inv_dt = 0, dt_ratio = 0; int velocity_iterations = 8, position_iterations = 3; bool warm_starting = true; };

// Per-contact input of the solver, detached from the world objects. This is what the island
// solver reads from a contact (src/private/dynamics/b2_contact_solver_private.rs:20-110).
struct SolverContact {
  Manifold* manifold;
  int indexA, indexB;
  real radiusA, radiusB;
  real friction, restitution, threshold, tangent_speed;
};
struct SolverBody { real inv_mass, inv_I; Vec2 local_center; };

struct ContactSolver {
  TimeStep step;
  std::vector<VelocityConstraint> vcs;
  std::vector<PositionConstraint> pcs;
  bool block_solve = true;
  void init(const std::vector<SolverContact>& contacts, const std::vector<SolverBody>& bodies);
  void initialize_velocity_constraints(const std::vector<SolverContact>& contacts, const Position* pos, const Velocity* vel);
  void warm_start(Velocity* vel);
  void solve_velocity_constraints(Velocity* vel);
  void store_impulses(const std::vector<SolverContact>& contacts);
  bool solve_position_constraints(Position* pos, real* min_sep_out = nullptr);
  bool solve_toi_position_constraints(Position* pos, int toi_index_a, int toi_index_b);
};

struct Profile { double step = 0, collide = 0, solve = 0, broadphase = 0; };

struct World {
  Vec2 gravity;
  std::vector<Body*> bodies;       // creation order; iterated in reverse = push_front list order
  std::vector<Fixture*> fixtures;  // creation order
  std::vector<Joint*> joints;      // creation order; iterated in reverse = push_front list order
  Contact* contact_list = nullptr;
  int contact_count = 0;
  BroadPhase bp;
  bool warm_starting = true, allow_sleep = true, continuous = false, block_solve = true;
  bool new_contacts = false;
  // single-stepped continuous physics: a TOI pass ends after its first solved event and the next
  // step resumes it without solving (src/b2_world.rs:188-193, src/private/dynamics/b2_world.rs:
  // 545, 678, 896-897, 935)
  bool sub_stepping = false, step_complete = true;
  real inv_dt0 = 0;
  int toi_events = 0;  // TOI events solved in the last step (a count for tests)
  Profile profile;

  explicit World(Vec2 g) : gravity(g) {}
  ~World();
  Body* create_body(int type, Vec2 pos, real angle, Vec2 v, real w, real lin_damp, real ang_damp, real gscale,
                    bool allow_sleep, bool awake, bool fixed_rotation);
  Fixture* create_fixture(Body* b, const Shape& s, real density, real friction, real restitution, real threshold,
                          Filter filter, bool sensor);
  void reset_mass_data(Body* b);
  Joint* create_joint(const Joint& def);
  void step(real dt, int vel_iters, int pos_iters);
  // contact manager: src/private/dynamics/b2_contact_manager.rs
  void collide();
  void find_new_contacts();
  void add_pair(int fixtureA, int fixtureB);
  void destroy_contact(Contact* c);
  void solve(const TimeStep& step);
  void solve_toi(const TimeStep& step);
  void update_contact(Contact* c);
  void synchronize_fixtures(Body* b);
};

// Island-level solve over an explicit body list and constraint order. This is the reference
// island solve (src/private/dynamics/b2_island_private.rs:129-328) with the sleep
// bookkeeping left to the caller; it is shared by the world step and the colour-order replay.
struct IslandIO {
  // bodies of the island, in island-index order
  std::vector<Body*> bodies;
  std::vector<Contact*> contacts;
  std::vector<Joint*> joints;
};
bool solve_island(World* w, IslandIO& island, const TimeStep& step);

}  // namespace b2o
