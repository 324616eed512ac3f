// b2g internals: world state, device buffers and the phase drivers shared by the .hip files.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/b2g.h"
#include "b2g_collide.h"

namespace b2g {

void set_error(const std::string& msg);
#define B2G_HIP(call)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));                          \
      return B2G_ERR_DEVICE;                                                                 \
    }                                                                                        \
  } while (0)

// A typed device array with capacity growth.
template <class T>
struct DevArray {
  T* p = nullptr;
  size_t cap = 0;
  int reserve(size_t n, bool keep = false, size_t keep_n = 0) {
    if (n <= cap) return B2G_OK;
    size_t ncap = cap ? cap : 64;
    while (ncap < n) ncap *= 2;
    T* q = nullptr;
    if (hipMalloc(&q, ncap * sizeof(T)) != hipSuccess) {
      set_error("hipMalloc failed");
      return B2G_ERR_DEVICE;
    }
    if (keep && p && keep_n)
      if (hipMemcpy(q, p, keep_n * sizeof(T), hipMemcpyDeviceToDevice) != hipSuccess) {
        set_error("hipMemcpy failed");
        return B2G_ERR_DEVICE;
      }
    if (p) (void)hipFree(p);
    p = q;
    cap = ncap;
    return B2G_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Host-side shape after construction (hull, box, ...); the device keeps it as a ShapeV record.
struct HostShape {
  int type = kCircle;
  int count = 0;
  int one_sided = 0;
  float radius = 0.0f;
  V2 p = {0.0f, 0.0f};       // circle centre / polygon centroid
  V2 v[8] = {};              // polygon vertices / edge v1, v2, v0, v3
  V2 n[8] = {};              // polygon normals
};
int build_host_shape(const b2g_shape_def& def, HostShape* out);  // B2G_OK or B2G_ERR_DEGENERATE
void host_shape_mass(const HostShape& s, float density, float* mass, V2* center, float* I);
ShapeV host_shape_view(const HostShape& s);

// Device-side body arrays (SoA, index = body creation order).
struct BodyArrays {
  DevArray<float4> pos;    // sweep c.x, c.y, a, 0
  DevArray<float4> vel;    // v.x, v.y, w, 0
  DevArray<float4> pos0;   // sweep c0.x, c0.y, a0, 0
  DevArray<float4> xf;     // p.x, p.y, sin, cos
  DevArray<float4> mass;   // inv_mass, inv_I, local_center.x, local_center.y
  DevArray<float4> props;  // mass, gravity_scale, linear_damping, angular_damping
  DevArray<float4> force;  // force.x, force.y, torque, 0
  DevArray<int> type;      // 0 static, 1 kinematic, 2 dynamic
  // sleep (used when the world can have sleeping bodies)
  DevArray<int> awake;       // 1 awake, 0 asleep (static bodies: 0)
  DevArray<float> sleep_time;
  DevArray<int> auto_sleep;  // the body def's allow_sleep
  DevArray<int> wake;        // wake request from the collide phase (touching changed, touching contact gone)
  DevArray<int> bullet;      // B2body's bullet flag: continuous collision also against dynamic bodies
};

// Device-side proxy arrays (one proxy per fixture; index = fixture creation order, which is
// also the order of the reference proxy ids).
struct ProxyArrays {
  DevArray<int> body;
  DevArray<int4> shape;     // type | count << 8 | one_sided << 16, 0, radius bits, 0
  DevArray<float4> mat;     // friction, restitution, restitution threshold, 0
  DevArray<int2> filter;    // category | mask << 16, group
  DevArray<float4> aabb;    // swept AABB (proxy.aabb)
  DevArray<float4> fat;     // tree fat AABB
  DevArray<int> moved;      // the tree's moved flag (set when the fat AABB was re-fattened)
  DevArray<int> world;      // world id of the proxy's body (batched worlds never interact)
  DevArray<int> tree_list;  // proxies kept in the BVH
  DevArray<int> loose_off;  // per world: range in `loose` of its big static proxies
  DevArray<int> loose;      // big static proxies (ground, walls) tested directly, not in the BVH
  DevArray<ShapeV> shapev;  // per proxy: its shape as the collide functions read it (by reference)
};

// A sensor contact has no manifold (point count 0, everything else zero, as the reference's
// never-evaluated manifold) and keeps its overlap state in this bit of the manifold header
// (src/private/dynamics/b2_contact.rs:148-162).
constexpr int kSensorTouching = 0x80;
constexpr int kSensorPending = 0x40;  // within the collide phase: a sensor contact to be tested

// Contact set: pairs grouped by their emitting proxy, plus manifolds.
struct ContactArrays {
  DevArray<int2> pair;       // (emitter proxy, other proxy)
  DevArray<int> off;         // CSR offsets by emitter proxy (np + 1)
  DevArray<int> old_index;   // index of the same pair in the previous set, or -1
  DevArray<int2> ab;         // fixtures A, B after registry ordering
  DevArray<int> hdr;         // manifold type | point_count << 8 (| kSensorTouching)
  DevArray<float4> geo;      // local normal, local point
  DevArray<float4> pt0, pt1; // local point, normal impulse, tangent impulse
  DevArray<uint2> ids;
  DevArray<int> colour;      // solve colour (launch index) of a touching contact, else -1
};

// Continuous collision state of one contact set: cached time of impact (as a fraction of the
// step), TOI count and flags (bit 0 cached, bit 1 disabled by a TOI event).
struct ToiState {
  DevArray<float> toi;
  DevArray<int> count, flags;
};

// Per-step solver arrays (indexed by constraint, colour-sorted).
// body-count bucket of a world for the per-world solve's threads per world: < 64, 64-255, >= 256
inline int world_bucket(int bodies) { return bodies < 64 ? 0 : bodies < 256 ? 1 : 2; }

struct SolverArrays {
  // per-world solve: constraints of each (colour, world), their exclusive scan (colour-major), the
  // scatter cursors, the worlds' body ranges and the launch order of the worlds
  DevArray<int> cw_count, cw_off, cw_cursor, world_body_off, world_order;
  // island solve: constraints per island root, the small islands' counts and their exclusive scan,
  // each constraint's island root
  DevArray<int> isl_count, isl_small, isl_start, croot;
  DevArray<int> touch_flag, touch_scan, touch_list;   // compaction of touching contacts
  DevArray<int> k_contact;   // contact index of constraint
  DevArray<int2> k_bodies;
  DevArray<int4> k_info;     // manifold type, point count, velocity point count, island
  DevArray<float4> vnf;      // normal.x, normal.y, friction, tangent_speed
  DevArray<float4> vm;       // mA, iA, mB, iB
  DevArray<float4> vr0, vr1; // rA.x, rA.y, rB.x, rB.y per point
  DevArray<float4> vms;      // normal_mass0, tangent_mass0, normal_mass1, tangent_mass1
  DevArray<float2> vbias;    // velocity bias per point
  DevArray<float4> K, NM;    // block matrix and its inverse (ex.x, ex.y, ey.x, ey.y)
  DevArray<float4> lam;      // normal0, tangent0, normal1, tangent1 impulses
  DevArray<float4> plp;      // position constraint local points 0, 1
  DevArray<float4> pgeo;     // local normal, local point
  DevArray<float4> plc;      // local centre A, local centre B
  DevArray<float2> prad;     // radius A, radius B
  // colouring
  DevArray<int> deg, adj_off, adj_cursor, adj;  // body -> constraints (dynamic bodies only)
  DevArray<int> color_state;                    // (round << 8) | colour, or -1
  DevArray<int> color_off, color_cursor;  // color_off: launch offsets, then per colour id
                                                       // its base offset and its launch index
  DevArray<int> prev_colour;                    // colour of the contact's predecessor, or -1
  DevArray<int> order;                          // colour-sorted constraint -> compacted index
  DevArray<int> inv_order;                      // compacted index -> colour-sorted constraint
  // islands
  DevArray<int> parent, isl_violated, isl_done;
  DevArray<int> isl_awake, isl_restless;  // per island root: has an awake body; a body is not ready to sleep
};

// One internal LBVH node as a walk sees it: both child boxes, the children (internal node
// index, or ~sorted position for a leaf), the last sorted leaf below each child and the world
// range below each child. 64 bytes, so a visited node costs one cache line.
struct alignas(16) BvhNode {
  float4 box[2];
  int4 link;    // {left, right, last leaf under left (the split), last leaf under right}
  int4 worlds;  // {left first, left last, right first, right last}
};

// The node a walk visits: a binary node's (up to) four grandchildren, so a walk descends two
// levels per visited node. 128 bytes: one cache line. Unused slots have an empty box and
// last = -1; a leaf in the place of a child fills one slot.
struct alignas(16) Bvh4Node {
  float4 box[4];
  int link[4];  // binary internal node index, or ~sorted position for a leaf
  int last[4];  // last sorted leaf under the slot
  int wlo[4], whi[4];  // world range under the slot
};

struct BroadphaseArrays {
  DevArray<unsigned> keys, keys_alt;
  DevArray<unsigned> vals, vals_alt;
  DevArray<unsigned> hist;       // radix histograms
  DevArray<float4> seg;          // segment tree over the sorted leaf boxes (level 0 = leaves)
  DevArray<BvhNode> nodes;       // LBVH internal nodes
  DevArray<Bvh4Node> nodes4;     // per internal node: its grandchildren (the walk's nodes)
  DevArray<int4> leafrec;        // per sorted leaf: proxy, body, filter bits, group | shape type | body type
  DevArray<int> count;           // pairs per proxy
  DevArray<int> slots;           // first pairs of each query, column-major (slot j of query k at j * n + k)
  DevArray<float4> bounds;       // scene bounds and flags
  DevArray<int> scan_tmp;        // block sums for scans
  // single-pass scan: per tile its published (epoch, flag, sum) word; [0] tile ticket, [1] error
  DevArray<unsigned long long> scan_state;
  DevArray<int> scan_ctl;
  unsigned scan_epoch = 0;
  bool scan_ctl_ready = false;
};

struct StepProfile {
  float collide_ms = 0, solve_ms = 0, broadphase_ms = 0, step_ms = 0;
};

// Event timing of every launch of one kernel (B2G_KERNEL_*): start/end event pairs are reused
// and harvested wherever the stream has been synchronised.
struct KernelTimer {
  int kernel = B2G_KERNEL_NONE;
  std::vector<hipEvent_t> ev;          // 2 per pending launch
  std::vector<long long> pending_items;
  double ms = 0.0;
  long long launches = 0, items = 0;
  long long aux = 0;  // a second count known only after the launch (pairs found by the traversal)
};

// Recorded forward run for the adjoint. A slot is one touching contact of one step, in that
// step's solve order; slots of all steps are concatenated.
struct TapeStep {
  int n = 0;                    // slots of this step
  long long off = 0;            // first slot
  std::vector<int> color_off;   // colour offsets (relative to off)
  float dt = 0.0f, dt_ratio = 0.0f;
  int vel_iters = 0, pos_iters = 0;
  bool warm = true, block = true;
};
struct Tape {
  bool active = false;
  bool has_prev = false;        // the previous step was recorded (carry sources exist)
  int every = 16;               // body-state checkpoint interval in steps
  std::vector<TapeStep> steps;
  long long total = 0;
  DevArray<int4> slot;          // bodyA, bodyB, fixtureA, fixtureB
  DevArray<int4> meta;          // type | count << 8, carry source of point 0 and 1 (slot*2 + point or -1), runs
  DevArray<uint2> ids;          // feature ids (checked when the manifold is recomputed)
  DevArray<float4> lam0;        // impulses as collide carried them (the source when carry is -1)
  DevArray<float4> lam_store;   // impulses as store_impulses left them in the manifold
  DevArray<float4> c_lam0;      // contact-indexed scratch of the current step
  DevArray<int2> c_carry;       // per point: previous-set contact * 2 + point, or -1
  DevArray<long long> c_slot[2];// slot of each contact of the two contact sets, or -1
  DevArray<float4> ck_pos, ck_vel;  // body state at the start of every checkpointed step
  int ck_count = 0;
  int nb = 0;
};

// Joints (distance, revolute): definitions, persistent impulses and per-step solver rows,
// indexed by joint creation order.
struct JointArrays {
  DevArray<int2> bodies;   // body A, body B
  DevArray<int4> type;     // B2G_JOINT_*, flags (bit0 collide_connected, bit1 limit, bit2 motor; gear:
                           // bit3 / bit4 joint 1 / 2 prismatic), gear: bodies C, D
  DevArray<float4> d0;     // local anchor A, local anchor B
  DevArray<float4> d1;     // distance: length, min, max, stiffness; revolute: reference, lower, upper, max torque
  DevArray<float4> d2;     // distance: damping; revolute: motor speed
  DevArray<float4> d3;     // the other kinds' fourth parameter record (see upload_joints)
  DevArray<float4> imp;    // distance: impulse, lower, upper; revolute: impulse.x, .y, motor, lower
  DevArray<float> imp2;    // revolute: upper impulse
  DevArray<float4> r;      // rA, rB of the step
  DevArray<float4> k;      // distance: u, mass, soft mass; revolute: K.ex.x, K.ex.y (= K.ey.x), K.ey.y, axial mass
  DevArray<float4> k2;     // distance: gamma, bias, current length; revolute: angle
  DevArray<float4> k3;     // wheel and gear: the rest of the step's solver data
  DevArray<int> order;     // joints in colour order
  DevArray<int> nc_off;    // per body: offsets into nc_list (nb + 1)
  DevArray<int> nc_list;   // per body: bodies it is joined to by a joint that forbids their collision
  DevArray<int> nc_flag;   // per body: 1 when its nc_list is not empty
};

struct World {
  // settings
  V2 gravity;
  // B2world::new (src/private/dynamics/b2_world.rs:37-43): warm starting, continuous physics and
  // sleeping on
  bool warm_starting = true, allow_sleep = true, continuous = true, block_solve = true;
  // single-stepped continuous physics (B2world::set_sub_stepping, src/b2_world.rs:188-193): a TOI
  // pass ends after its first solved event and step_complete stays false until a pass finds no
  // event; a step with step_complete false skips the solve and resumes the pass without resetting
  // it (src/private/dynamics/b2_world.rs:545, 678, 896-897, 935)
  bool sub_stepping = false, step_complete = true;
  int n_sensors = 0;  // sensor fixtures (B2fixtureDef::is_sensor)
  int device = 0;  // the HIP device current at creation: the stream's device
  // sleep_possible: allow_sleep is on or a body was created asleep; otherwise every non-static body
  // is awake at every step and the sleep bookkeeping is skipped
  bool created_asleep = false;
  bool sleep_possible() const { return allow_sleep || created_asleep; }
  float inv_dt0 = 0.0f;
  // host mirrors of static data
  std::vector<b2g_body_def> body_defs;
  std::vector<int> host_type;
  std::vector<float4> h_pos, h_pos0, h_vel, h_xf, h_mass, h_props;
  std::vector<float> h_I;
  std::vector<std::vector<int>> body_fixtures;  // creation order
  std::vector<int> h_fbody;
  std::vector<float> h_fdensity;
  std::vector<int4> h_fshape;
  std::vector<float4> h_fmat, h_ffat, h_faabb;
  std::vector<int2> h_ffilter;
  std::vector<int> h_bworld, h_fworld;  // world id per body / per proxy
  int n_tree = 0;                       // proxies in the BVH
  // host sources of asynchronous uploads must outlive the copy: kept here, not on the stack
  std::vector<int> h_tree, h_loose_off, h_loose;
  int n_worlds = 1;
  std::vector<ShapeV> h_shapev;
  std::vector<int> h_awake, h_auto_sleep, h_bullet;
  std::vector<float> h_sleep_time;
  std::vector<HostShape> h_shapes;
  int nb = 0, np = 0;          // bodies, proxies on the host
  int dev_nb = 0, dev_np = 0;  // uploaded
  bool need_broadphase = false;  // new fixtures: find pairs before the next collide
  bool collide_pending = false;  // the current set has not been through a collide yet
  bool dirty = false;            // host items not yet on the device
  bool manifolds_valid[2] = {true, true};  // set went through a narrowphase (or a carry)
  bool profile_pending = false;
  // device state
  hipStream_t stream = nullptr;
  BodyArrays b;
  ProxyArrays f;
  ContactArrays c[2];  // current and previous contact sets
  ToiState toi_state[2];              // per contact set, when continuous physics is on
  // touching contacts a TOI rebuild drops from the predecessor role (flags, scan, and the
  // gathered entries appended to the new predecessor set; see toi_rebuild)
  DevArray<int> toi_succ, toi_drop_flag, toi_drop_pos, toi_drop_total;
  ContactArrays toi_dropped;
  DevArray<float> toi_alpha0;         // per body: the sweep's alpha0 during the TOI pass
  DevArray<unsigned char> toi_big;    // TOI islands too large for the event workgroup's LDS
  bool toi_force_big = false;         // every TOI island in toi_big (a check of that path)
  DevArray<unsigned long long> toi_ctl;  // smallest TOI key, event status, event bodies, counters
  DevArray<int> toi_cand, toi_need, toi_rank;  // TOI event candidates; contacts to recompute, ranked
  DevArray<float> toi_snap;                    // aligned sweeps of the contacts to recompute
  // device event loop: contacts per non-static body of the current set, fixtures per body (built
  // when the fixture count changes), the event candidates (contacts with a cached time of impact
  // below 1) and their membership flags
  DevArray<int> toi_cadj_off, toi_cadj, toi_cadj_cur, toi_bfix_off, toi_bfix, toi_bfix_cur;
  DevArray<int> toi_cl, toi_in_list;
  int toi_bfix_np = -1;
  int toi_events = 0;                 // TOI events solved in the last step
  int cur = 0;         // index of the current contact set in c[]
  int n_pairs = 0;     // pairs in the current set
  int n_pairs_old = 0; // pairs in the previous set
  int set_np[2] = {0, 0};  // proxy count when each contact set was built (0: never built)
  SolverArrays s;
  BroadphaseArrays bp;
  DevArray<int> counters;   // small device counters: 4 solver counters, then the colour histogram
  // pinned host memory (kPinnedInts ints): [0, 134) the solver's counters read back after the
  // colouring (4 counters, 128 colour counts, last round, large-island constraints),
  // [kPinnedPairFlags, +4) the pair search's flags and total, [kPinnedColourOffsets, +385) the
  // colour offsets uploaded for the solve
  int* h_counters = nullptr;
  int n_touch = 0, n_colors = 0, n_islands = 0;
  int color_rounds = 16;       // colouring rounds launched before the first readback
  bool color_reuse = true;     // persisting contacts keep their previous colour
  int scratch_colors = 0;      // colours of the last from-scratch colouring (0: none yet)
  int reuse_rounds = 0;        // colouring rounds the last reusing step needed
  int steps_since_scratch = 0; // solves since the last from-scratch colouring
  bool solver_graphs = false;  // replay the solve's per-colour launches as a captured HIP graph
  // Batched worlds: each world's velocity and position iterations run in one workgroup per world
  // (colour after colour with workgroup barriers) instead of one launch per colour and iteration.
  // world_solve: -1 auto (on when the batch has at least kWorldSolveMin worlds), 0 off, 1 on.
  // world_threads[k]: threads per world for the worlds of body-count bucket k (see world_bucket).
  int world_solve = -1;
  int world_threads[3] = {64, 128, 256};
  // island_solve: -1 auto (on unless the per-world solve is), 0 off, 1 on; bin_threads: threads per
  // bin (islands of at most bin_threads / 2 constraints are binned)
  int island_solve = -1;
  int bin_threads = 256;
  int n_bins = 0;  // bins of the last solve
  bool worlds_contiguous = false;         // every world's bodies are one index range (set at upload)
  std::vector<int> h_world_body_off;      // n_worlds + 1 (when contiguous)
  std::vector<int> h_world_order;         // worlds by bucket, larger worlds first within a bucket
  int world_bucket_off[4] = {0, 0, 0, 0}; // ranges of h_world_order per bucket
  struct CachedGraph {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
    unsigned long long last_use = 0;
  };
  std::vector<CachedGraph> graphs;
  unsigned long long graph_clock = 0;
  int color_rounds_used = 0;
  std::vector<int> h_color_off;
  int pos_iters_run = 0;
  StepProfile prof;
  hipEvent_t ev[8] = {};
  KernelTimer kt;
  Tape tape;
  // joints
  int nj = 0, dev_nj = 0;
  int n_nocollide = 0;  // entries of the per-body no-collide lists
  std::vector<b2g_joint_def> h_joints;
  std::vector<float> h_joint_constant;  // per joint: a pulley's or a gear's constant
  std::vector<int> h_joint_order, h_joint_colour_off;
  bool joints_dirty = false;
  JointArrays j;
};

constexpr int kPinnedInts = 1024, kPinnedPairFlags = 192, kPinnedColourOffsets = 256;

// Capacity floor of the arrays sized by the contact set (the pairs, and the solver's arrays,
// which the pairs bound): reserving for a multiple of the proxy count from the first step keeps a
// contact set that grows while a scene settles (the 65,536-body mixed world goes from about one
// to about 6.5 pairs per proxy) from reallocating step after step, each hipFree waiting for the
// device. Beyond the floor the arrays still double as needed.
constexpr size_t kPairsPerProxy = 8;
inline size_t pair_capacity(const World* w, size_t n) {
  return std::max(std::max(n, size_t(1)), size_t(w->np) * kPairsPerProxy);
}

// adjoint tape hooks of b2g_world_step (b2g_adjoint.hip)
int tape_step_begin(World* w);
int tape_after_collide(World* w);
int tape_after_solve(World* w, float dt, float dt_ratio, int vel_iters, int pos_iters);
void tape_release(World* w);
int tape_replay(World* w, float* final_state);
int tape_backward(World* w, const float* cotangent, float* grad_state, float* grad_friction);

// Launches a kernel; when it is the timed kernel of the world, the launch carries a start/stop
// event pair (hipExtLaunchKernelGGL records them around the kernel's own dispatch, so the
// elapsed time is the kernel's run, not the queue ahead of it). Pairs are harvested wherever the
// stream has been synchronised.
template <typename... Args, typename F>
inline void launch_timed(World* w, int kernel, long long items, F func, dim3 grid, dim3 block, Args... args) {
  KernelTimer& t = w->kt;
  if (t.kernel == kernel) {
    const size_t k = 2 * t.pending_items.size();
    bool ok = true;
    while (ok && t.ev.size() < k + 2) {
      hipEvent_t e;
      ok = hipEventCreate(&e) == hipSuccess;
      if (ok) t.ev.push_back(e);
    }
    if (ok) {
      hipExtLaunchKernelGGL(func, grid, block, 0, w->stream, t.ev[k], t.ev[k + 1], 0, args...);
      t.pending_items.push_back(items);
      return;
    }
  }
  hipLaunchKernelGGL(func, grid, block, 0, w->stream, args...);
}
inline void kt_aux(World* w, int kernel, long long value) {
  if (w->kt.kernel == kernel) w->kt.aux += value;
}
int kt_harvest(World* w);

// phase drivers (each returns a B2G_* status)
int upload_world(World* w);
int phase_broadphase(World* w);  // fat AABBs -> next contact set (pairs + old indices)
int phase_collide(World* w);     // narrowphase on the current contact set
int phase_solve(World* w, float dt, float dt_ratio, int vel_iters, int pos_iters);
int phase_synchronize(World* w); // swept + fat AABBs of non-static proxies
int phase_toi(World* w, float dt, int vi);  // continuous collision after the solve
int toi_carry_state(World* w);              // TOI state of the previous contact set onto the current one
void release_solve_graphs(World* w);
// primitives
int exclusive_scan(World* w, const int* in, int* out, int n, int* total_dev, int* total2 = nullptr);
int radix_sort_pairs(World* w, unsigned* keys, unsigned* vals, int n, int bits);
// stateless collide of fixture pairs under given transforms (parity entry point)
int collide_fixture_pairs(World* w, int n, const int32_t* pairs, const float* xfa, const float* xfb,
                          b2g_manifold* out);
int collide_fixture_pairs_noalloc(World* w, int n, const int32_t* pairs, const float* xfa, const float* xfb,
                                  b2g_manifold* out);
// looks up the address of the no-allocation scratch array on `device` (which must be current)
int noalloc_scratch_init(int device);
int broadphase_pairs_for(World* w, const float* fat, int64_t cap, int32_t* out, int64_t* count);

// launch geometry helpers
inline int blocks_for(long long n, int bs) { return int((n + bs - 1) / bs); }

// float <-> order-preserving int for atomicMin on floats
__host__ __device__ inline int float_flip(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7fffffff;
}
__host__ __device__ inline float float_unflip(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

// Which bodies this step's solve left alone: an island (connected component of the touching
// contacts through non-static bodies) is solved only when one of its bodies is awake.
struct IslandView {
  const int* parent;     // body -> island root
  const int* isl_awake;  // per root; null: every island is awake
  __device__ bool asleep(int b) const { return isl_awake && !isl_awake[parent[b]]; }
};
// phase_synchronize for the bodies the view holds awake; any_moved (optional) is set to 1 when a
// proxy is re-fattened
int phase_synchronize_view(World* w, IslandView iv, int* any_moved);

// Kernel arguments of the joint solve (b2g_joints.hip).
struct JointArgs {
  const int* order;
  const int2* jb;
  const int4* jt;
  const float4 *jd0, *jd1, *jd2, *jd3;
  float4* jimp;
  float* jimp2;
  float4 *jr, *jk, *jk2, *jk3;
  const int* btype;
  float4* pos;
  float4* vel;
  const float4* mass;
  IslandView iv;
  int* violated;
  const int* done;
  float dt, inv_dt, dt_ratio;
  int warm;
};

// joint solve (b2g_joints.hip): phase 0 init + warm start, 1 velocity, 2 position; one launch
// per joint colour
void launch_joint_colours(World* w, int phase, const JointArgs& a);
int upload_joints(World* w);
// joints join islands (b2g_solver.hip)
void launch_joint_union(World* w);


// body transform record (p.x, p.y, sin, cos) -> Xf
__device__ inline Xf xf_of(float4 t) {
  Xf x;
  x.p = v2(t.x, t.y);
  x.q.s = t.z;
  x.q.c = t.w;
  return x;
}

// B2fixture::synchronize + B2dynamicTree::move_proxy for proxy i of body b
// (src/private/dynamics/b2_fixture.rs:147-173, src/private/collision/b2_dynamic_tree.rs:109-168):
// the swept AABB from the sweep's start (pos0) to the current transform, and the fat AABB
// re-fattened (with the displacement prediction) when it no longer contains the swept one or has
// grown too large; returns whether it was re-fattened.
__device__ inline bool sync_proxy(int i, int b, const ShapeV& s, const float4* __restrict__ pos0,
                                  const float4* __restrict__ bmass, const float4* __restrict__ bxf,
                                  float4* __restrict__ faabb, float4* __restrict__ ffat) {
  const float4 p0 = pos0[b];
  const float4 ms = bmass[b];
  Xf xf1;
  xf1.q = rot(p0.z);
  xf1.p = v2(p0.x, p0.y) - mul(xf1.q, v2(ms.z, ms.w));
  const Xf xf2 = xf_of(bxf[b]);
  const Aabb a1 = shape_aabb(s, xf1);
  const Aabb a2 = shape_aabb(s, xf2);
  const Aabb a = combine(a1, a2);
  faabb[i] = make_float4(a.lo.x, a.lo.y, a.hi.x, a.hi.y);
  const V2 disp = center(a2) - center(a1);
  const V2 r = v2(kAabbExtension, kAabbExtension);
  Aabb fat;
  fat.lo = a.lo - r;
  fat.hi = a.hi + r;
  const V2 d = kAabbMultiplier * disp;
  if (d.x < 0.0f) fat.lo.x += d.x; else fat.hi.x += d.x;
  if (d.y < 0.0f) fat.lo.y += d.y; else fat.hi.y += d.y;
  const float4 t = ffat[i];
  Aabb tree;
  tree.lo = v2(t.x, t.y);
  tree.hi = v2(t.z, t.w);
  if (contains(tree, a)) {
    Aabb huge;
    huge.lo = fat.lo - 4.0f * r;
    huge.hi = fat.hi + 4.0f * r;
    if (contains(huge, tree)) return false;
  }
  ffat[i] = make_float4(fat.lo.x, fat.lo.y, fat.hi.x, fat.hi.y);
  return true;
}

}  // namespace b2g
