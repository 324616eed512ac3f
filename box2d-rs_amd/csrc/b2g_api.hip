// b2g C ABI: world lifecycle, host-side creation of bodies and fixtures, upload, the step
// driver and readbacks. See include/b2g.h for the contract of each entry point.
#include <algorithm>
#include <cstring>

#include <unordered_map>

#include "b2g_internal.h"

struct b2g_world {
  b2g::World w;
};

namespace b2g {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }

static constexpr int kBlock = 256;

// carry manifolds of the previous set into the current one through old_index (new pairs get
// empty manifolds): used when the current set has not been through a narrowphase yet
__global__ void __launch_bounds__(kBlock) k_carry(int n, const int* __restrict__ old_index, const int2* __restrict__ pairs,
                                                  const int4* __restrict__ fshape, const int* __restrict__ ohdr,
                                                  const float4* __restrict__ ogeo, const float4* __restrict__ opt0,
                                                  const float4* __restrict__ opt1, const uint2* __restrict__ oids,
                                                  const int2* __restrict__ oab, int2* __restrict__ ab, int* __restrict__ hdr,
                                                  float4* __restrict__ geo, float4* __restrict__ pt0,
                                                  float4* __restrict__ pt1, uint2* __restrict__ ids,
                                                  int* __restrict__ touching) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int o = old_index[k];
  if (o >= 0) {
    ab[k] = oab[o];
    hdr[k] = ohdr[o];
    geo[k] = ogeo[o];
    pt0[k] = opt0[o];
    pt1[k] = opt1[o];
    ids[k] = oids[o];
    touching[k] = (ohdr[o] >> 8) > 0;
  } else {
    int2 pr = pairs[k];
    int lo = pr.x < pr.y ? pr.x : pr.y, hi = pr.x < pr.y ? pr.y : pr.x;
    bool swap = registry_swap(fshape[lo].x & 0xff, fshape[hi].x & 0xff) == 1;
    ab[k] = swap ? make_int2(hi, lo) : make_int2(lo, hi);
    hdr[k] = 0;
    geo[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    pt0[k] = geo[k];
    pt1[k] = geo[k];
    ids[k] = make_uint2(0u, 0u);
    touching[k] = 0;
  }
}

__global__ void k_wake_all(int nb, const int* __restrict__ btype, int* __restrict__ awake, float* __restrict__ sleep_time) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb || btype[b] == 0) return;
  awake[b] = 1;
  sleep_time[b] = 0.0f;
}

static int carry_manifolds(World* w) {
  if (w->manifolds_valid[w->cur]) return B2G_OK;
  int n = w->n_pairs;
  int st;
  if ((st = w->s.touch_flag.reserve(pair_capacity(w, n)))) return st;
  ContactArrays& c = w->c[w->cur];
  ContactArrays& o = w->c[1 - w->cur];
  if (n > 0) {
    hipLaunchKernelGGL(k_carry, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, w->stream, n, c.old_index.p, c.pair.p,
                       w->f.shape.p, o.hdr.p, o.geo.p, o.pt0.p, o.pt1.p, o.ids.p, o.ab.p, c.ab.p, c.hdr.p, c.geo.p,
                       c.pt0.p, c.pt1.p, c.ids.p, w->s.touch_flag.p);
    B2G_HIP(hipGetLastError());
  }
  w->manifolds_valid[w->cur] = true;
  return B2G_OK;
}

int carry_manifolds_for_toi(World* w) { return carry_manifolds(w); }

// ---- host-side creation (create_body / create_fixture / reset_mass_data) --------------------
// src/private/dynamics/b2_body.rs:15-89 (body), :292-349 (reset_mass_data)
static void reset_mass(World* w, int b) {
  float4& ms = w->h_mass[b];
  float4& pr = w->h_props[b];
  float4& pos = w->h_pos[b];
  float4& vel = w->h_vel[b];
  const float4 xf = w->h_xf[b];
  float mass = 0.0f, I = 0.0f, inv_mass = 0.0f, inv_I = 0.0f;
  V2 lc = v2(0.0f, 0.0f);
  int type = w->host_type[b];
  if (type == B2G_STATIC || type == B2G_KINEMATIC) {
    ms = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    pr.x = 0.0f;
    pos.x = xf.x;
    pos.y = xf.y;
    w->h_pos0[b] = pos;
    return;
  }
  // fixtures of the body, newest first (the reference's fixture list order)
  const std::vector<int>& fl = w->body_fixtures[b];
  for (auto it = fl.rbegin(); it != fl.rend(); ++it) {
    int f = *it;
    float density = w->h_fdensity[f];
    if (density == 0.0f) continue;
    float m;
    V2 c;
    float i;
    host_shape_mass(w->h_shapes[f], density, &m, &c, &i);
    mass += m;
    lc = lc + m * c;
    I += i;
  }
  if (mass > 0.0f) {
    inv_mass = 1.0f / mass;
    lc = inv_mass * lc;
  }
  if (I > 0.0f && !w->body_defs[b].fixed_rotation) {
    I -= mass * dot(lc, lc);
    inv_I = 1.0f / I;
  } else {
    I = 0.0f;
    inv_I = 0.0f;
  }
  V2 old_c = v2(pos.x, pos.y);
  Xf t;
  t.p = v2(xf.x, xf.y);
  t.q.s = xf.z;
  t.q.c = xf.w;
  V2 c = mul(t, lc);
  pos.x = c.x;
  pos.y = c.y;
  w->h_pos0[b] = pos;
  V2 dv = cross(vel.z, c - old_c);
  vel.x += dv.x;
  vel.y += dv.y;
  ms = make_float4(inv_mass, inv_I, lc.x, lc.y);
  pr.x = mass;
  w->h_I[b] = I;
}

static int add_items(World* w, int nbn, const b2g_body_def* bodies, const int32_t* body_world, int nf,
                     const int32_t* fixture_body, const b2g_shape_def* shapes, const b2g_fixture_def* fixtures,
                     int32_t* proxy_ids) {
  if (!w->step_complete && (nbn > 0 || nf > 0)) {
    set_error("bodies and fixtures cannot be added while a sub-stepped TOI pass is unfinished");
    return B2G_ERR_UNSUPPORTED;
  }
  for (int i = 0; i < nbn; ++i) {
    const b2g_body_def& d = bodies[i];
    if (body_world && (body_world[i] < 0 || body_world[i] >= (1 << 20))) {
      set_error("world id out of range [0, 2^20)");
      return B2G_ERR_INVALID_ARGUMENT;
    }
    if (d.type < 0 || d.type > 2) { set_error("invalid body type"); return B2G_ERR_INVALID_ARGUMENT; }
    if (d.linear_damping < 0.0f || d.angular_damping < 0.0f) {
      set_error("damping must be non-negative");
      return B2G_ERR_INVALID_ARGUMENT;
    }
  }
  const int base = w->nb;
  for (int k = 0; k < nf; ++k) {
    // relative to the first body of this call; negative values address earlier bodies
    int b = fixture_body[k];
    if (base + b < 0 || b >= nbn) {
      set_error("fixture_body index out of range");
      return B2G_ERR_INVALID_ARGUMENT;
    }
  }
  // build all shapes first so that a degenerate polygon leaves the world unchanged
  std::vector<HostShape> hs(nf);
  for (int k = 0; k < nf; ++k) {
    int st = build_host_shape(shapes[k], &hs[k]);
    if (st != B2G_OK) {
      set_error(st == B2G_ERR_DEGENERATE ? "degenerate polygon" : "unsupported shape type");
      return st;
    }
  }
  for (int i = 0; i < nbn; ++i) {
    const b2g_body_def& d = bodies[i];
    w->body_defs.push_back(d);
    w->host_type.push_back(d.type);
    Rot q = rot(d.angle);
    w->h_xf.push_back(make_float4(d.position[0], d.position[1], q.s, q.c));
    float4 p = make_float4(d.position[0], d.position[1], d.angle, 0.0f);
    w->h_pos.push_back(p);
    w->h_pos0.push_back(p);
    w->h_vel.push_back(make_float4(d.linear_velocity[0], d.linear_velocity[1], d.angular_velocity, 0.0f));
    // B2body::new: awake unless the def says otherwise; static bodies are never awake
    const int awake = d.type != B2G_STATIC && d.awake ? 1 : 0;
    if (d.type != B2G_STATIC && !awake) w->created_asleep = true;
    w->h_awake.push_back(awake);
    w->h_auto_sleep.push_back(d.allow_sleep ? 1 : 0);
    w->h_bullet.push_back(d.bullet ? 1 : 0);
    w->h_sleep_time.push_back(0.0f);
    w->h_mass.push_back(make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    w->h_props.push_back(make_float4(0.0f, d.gravity_scale, d.linear_damping, d.angular_damping));
    w->h_I.push_back(0.0f);
    w->body_fixtures.emplace_back();
    int wid = body_world ? body_world[i] : 0;
    w->h_bworld.push_back(wid);
    w->n_worlds = std::max(w->n_worlds, wid + 1);
  }
  w->nb += nbn;
  for (int k = 0; k < nf; ++k) {
    int b = base + fixture_body[k];
    const b2g_fixture_def& fd = fixtures[k];
    const HostShape& s = hs[k];
    int f = w->np;
    w->h_shapes.push_back(s);
    w->h_fbody.push_back(b);
    w->h_fdensity.push_back(fd.density);
    {
      // the shape as the device kernels read it (fields a type does not use stay zero)
      ShapeV r;
      std::memset(&r, 0, sizeof(r));
      r.type = s.type;
      r.count = s.count;
      r.one_sided = s.one_sided;
      r.radius = s.radius;
      if (s.type == kCircle) {
        r.p = s.p;
      } else if (s.type == kPolygon) {
        for (int i = 0; i < s.count; ++i) { r.v[i] = s.v[i]; r.n[i] = s.n[i]; }
        r.p = s.p;
      } else {
        for (int i = 0; i < 4; ++i) r.v[i] = s.v[i];
      }
      w->h_shapev.push_back(r);
    }
    int radius_bits;
    std::memcpy(&radius_bits, &s.radius, 4);
    // .w: the fixture is a sensor (B2fixtureDef::is_sensor)
    if (fixtures[k].is_sensor) ++w->n_sensors;
    w->h_fshape.push_back(
        make_int4(s.type | (s.count << 8) | (s.one_sided << 16), 0, radius_bits, fixtures[k].is_sensor ? 1 : 0));
    w->h_fmat.push_back(make_float4(fd.friction, fd.restitution, fd.restitution_threshold, 0.0f));
    w->h_ffilter.push_back(make_int2(int(fd.category_bits) | (int(fd.mask_bits) << 16), fd.group_index));
    w->h_fworld.push_back(w->h_bworld[b]);
    // create_proxies: AABB at the body transform, fattened by the AABB extension
    // (src/private/dynamics/b2_fixture.rs:114-133, src/private/collision/b2_dynamic_tree.rs:81-99)
    float4 x = w->h_xf[b];
    Xf t;
    t.p = v2(x.x, x.y);
    t.q.s = x.z;
    t.q.c = x.w;
    Aabb a = shape_aabb(host_shape_view(s), t);
    w->h_faabb.push_back(make_float4(a.lo.x, a.lo.y, a.hi.x, a.hi.y));
    w->h_ffat.push_back(make_float4(a.lo.x - kAabbExtension, a.lo.y - kAabbExtension, a.hi.x + kAabbExtension,
                                    a.hi.y + kAabbExtension));
    w->body_fixtures[b].push_back(f);
    w->np += 1;
    if (fd.density > 0.0f) reset_mass(w, b);
    // proxy id the reference's dynamic tree assigns in a create-only world: node 0 for the
    // first leaf, then each insert allocates a parent and the leaf (2k - 1)
    if (proxy_ids) proxy_ids[k] = f == 0 ? 0 : 2 * f - 1;
  }
  w->dirty = true;
  w->need_broadphase = true;
  return B2G_OK;
}

// ---- upload / download ---------------------------------------------------------------------
template <class T>
static int up(DevArray<T>& d, const std::vector<T>& h, hipStream_t s) {
  int st = d.reserve(std::max<size_t>(h.size(), 1));
  if (st) return st;
  if (!h.empty()) B2G_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
  return B2G_OK;
}
template <class T>
static int down(std::vector<T>& h, const DevArray<T>& d, size_t n, hipStream_t s) {
  h.resize(n);
  if (n) B2G_HIP(hipMemcpyAsync(h.data(), d.p, sizeof(T) * n, hipMemcpyDeviceToHost, s));
  return B2G_OK;
}

// pulls the device-owned dynamic state into the host mirrors
static int download_state(World* w) {
  if (w->dev_nb == 0 && w->dev_np == 0) return B2G_OK;
  std::vector<float4> pos, vel, pos0, xf, fat, faabb;
  std::vector<int> awake;
  std::vector<float> sleep_time;
  int st;
  if ((st = down(pos, w->b.pos, w->dev_nb, w->stream)) || (st = down(vel, w->b.vel, w->dev_nb, w->stream)) ||
      (st = down(pos0, w->b.pos0, w->dev_nb, w->stream)) || (st = down(xf, w->b.xf, w->dev_nb, w->stream)) ||
      (st = down(fat, w->f.fat, w->dev_np, w->stream)) || (st = down(faabb, w->f.aabb, w->dev_np, w->stream)) ||
      (st = down(awake, w->b.awake, w->dev_nb, w->stream)) ||
      (st = down(sleep_time, w->b.sleep_time, w->dev_nb, w->stream)))
    return st;
  B2G_HIP(hipStreamSynchronize(w->stream));
  std::copy(awake.begin(), awake.end(), w->h_awake.begin());
  std::copy(sleep_time.begin(), sleep_time.end(), w->h_sleep_time.begin());
  std::copy(pos.begin(), pos.end(), w->h_pos.begin());
  std::copy(vel.begin(), vel.end(), w->h_vel.begin());
  std::copy(pos0.begin(), pos0.end(), w->h_pos0.begin());
  std::copy(xf.begin(), xf.end(), w->h_xf.begin());
  std::copy(fat.begin(), fat.end(), w->h_ffat.begin());
  std::copy(faabb.begin(), faabb.end(), w->h_faabb.begin());
  return B2G_OK;
}

static bool host_is_static(const World* w, int f) { return w->host_type[w->h_fbody[f]] == B2G_STATIC; }

// Joint rows, the joint colouring and the per-body lists of bodies a joint keeps from colliding.
// Impulses of joints already on the device are kept; new joints start at zero.
int upload_joints(World* w) {
  const int nj = w->nj, nb = w->nb;
  int st;
  std::vector<int2> jb(nj);
  std::vector<int4> jt(nj);
  std::vector<float4> d0(nj), d1(nj), d2(nj), d3(nj, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  std::vector<std::vector<int>> nc(nb);
  for (int i = 0; i < nj; ++i) {
    const b2g_joint_def& d = w->h_joints[i];
    jb[i] = make_int2(d.body_a, d.body_b);
    jt[i] = make_int4(d.type, (d.collide_connected ? 1 : 0) | (d.enable_limit ? 2 : 0) | (d.enable_motor ? 4 : 0), 0, 0);
    d0[i] = make_float4(d.local_anchor_a[0], d.local_anchor_a[1], d.local_anchor_b[0], d.local_anchor_b[1]);
    if (d.type == B2G_JOINT_DISTANCE) {
      d1[i] = make_float4(d.length, d.min_length, d.max_length, d.stiffness);
      d2[i] = make_float4(d.damping, 0.0f, 0.0f, 0.0f);
    } else if (d.type == B2G_JOINT_WELD) {
      d1[i] = make_float4(d.reference_angle, d.stiffness, d.damping, 0.0f);
      d2[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else if (d.type == B2G_JOINT_PRISMATIC) {
      // the axis normalised as B2vec2::normalize does (src/b2_math.rs:82-92)
      V2 axis = v2(d.local_axis_a[0], d.local_axis_a[1]);
      normalize(&axis);
      d1[i] = make_float4(d.reference_angle, d.lower_angle, d.upper_angle, d.max_motor_torque);
      d2[i] = make_float4(d.motor_speed, axis.x, axis.y, 0.0f);
    } else if (d.type == B2G_JOINT_FRICTION) {
      d1[i] = make_float4(d.max_force, d.max_torque, 0.0f, 0.0f);
      d2[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else if (d.type == B2G_JOINT_MOTOR) {
      d0[i] = make_float4(d.linear_offset[0], d.linear_offset[1], 0.0f, 0.0f);
      d1[i] = make_float4(d.max_force, d.max_torque, d.angular_offset, d.correction_factor);
      d2[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else if (d.type == B2G_JOINT_MOUSE) {
      d1[i] = make_float4(d.target[0], d.target[1], d.max_force, d.stiffness);
      d2[i] = make_float4(d.damping, 0.0f, 0.0f, 0.0f);
    } else if (d.type == B2G_JOINT_WHEEL) {
      // the axis as given (B2wheelJoint::new does not normalise it, src/joints/b2_wheel_joint.rs:334)
      d1[i] = make_float4(d.lower_angle, d.upper_angle, d.max_motor_torque, d.motor_speed);
      d2[i] = make_float4(d.stiffness, d.damping, d.local_axis_a[0], d.local_axis_a[1]);
    } else if (d.type == B2G_JOINT_PULLEY) {
      d1[i] = make_float4(d.ground_anchor_a[0], d.ground_anchor_a[1], d.ground_anchor_b[0], d.ground_anchor_b[1]);
      d2[i] = make_float4(d.ratio, w->h_joint_constant[i], 0.0f, 0.0f);
    } else if (d.type == B2G_JOINT_GEAR) {
      // bodies A and B as the coupled joints' bodies B, C and D as their bodies A, with their
      // anchors, axes (prismatic) and reference angles (src/private/dynamics/joints/b2_gear_joint.rs:8-85)
      const b2g_joint_def& j1 = w->h_joints[d.joint1];
      const b2g_joint_def& j2 = w->h_joints[d.joint2];
      V2 axis_c = v2(0.0f, 0.0f), axis_d = v2(0.0f, 0.0f);
      if (j1.type == B2G_JOINT_PRISMATIC) {
        axis_c = v2(j1.local_axis_a[0], j1.local_axis_a[1]);
        normalize(&axis_c);
      }
      if (j2.type == B2G_JOINT_PRISMATIC) {
        axis_d = v2(j2.local_axis_a[0], j2.local_axis_a[1]);
        normalize(&axis_d);
      }
      jt[i].y |= (j1.type == B2G_JOINT_PRISMATIC ? 8 : 0) | (j2.type == B2G_JOINT_PRISMATIC ? 16 : 0);
      jt[i].z = j1.body_a;
      jt[i].w = j2.body_a;
      d0[i] = make_float4(j1.local_anchor_b[0], j1.local_anchor_b[1], j2.local_anchor_b[0], j2.local_anchor_b[1]);
      d1[i] = make_float4(j1.local_anchor_a[0], j1.local_anchor_a[1], j2.local_anchor_a[0], j2.local_anchor_a[1]);
      d2[i] = make_float4(axis_c.x, axis_c.y, axis_d.x, axis_d.y);
      d3[i] = make_float4(j1.reference_angle, j2.reference_angle, d.ratio, w->h_joint_constant[i]);
    } else {
      d1[i] = make_float4(d.reference_angle, d.lower_angle, d.upper_angle, d.max_motor_torque);
      d2[i] = make_float4(d.motor_speed, 0.0f, 0.0f, 0.0f);
    }
    if (!d.collide_connected) {
      nc[d.body_a].push_back(d.body_b);
      nc[d.body_b].push_back(d.body_a);
    }
  }
  std::vector<int> nc_off(nb + 1, 0), nc_list, nc_flag(nb, 0);
  for (int b = 0; b < nb; ++b) {
    nc_off[b + 1] = nc_off[b] + int(nc[b].size());
    nc_flag[b] = nc[b].empty() ? 0 : 1;
    nc_list.insert(nc_list.end(), nc[b].begin(), nc[b].end());
  }
  JointArrays& J = w->j;
  // impulses: grow keeping the uploaded prefix
  const int had = w->dev_nj;
  if ((st = J.imp.reserve(std::max(nj, 1), true, had)) || (st = J.imp2.reserve(std::max(nj, 1), true, had)) ||
      (st = J.r.reserve(std::max(nj, 1))) || (st = J.k.reserve(std::max(nj, 1))) || (st = J.k2.reserve(std::max(nj, 1))) ||
      (st = J.k3.reserve(std::max(nj, 1))))
    return st;
  if (nj > had) {
    B2G_HIP(hipMemsetAsync(J.imp.p + had, 0, sizeof(float4) * (nj - had), w->stream));
    B2G_HIP(hipMemsetAsync(J.imp2.p + had, 0, sizeof(float) * (nj - had), w->stream));
  }
  if ((st = up(J.bodies, jb, w->stream)) || (st = up(J.type, jt, w->stream)) || (st = up(J.d0, d0, w->stream)) ||
      (st = up(J.d1, d1, w->stream)) || (st = up(J.d2, d2, w->stream)) || (st = up(J.d3, d3, w->stream)) ||
      (st = up(J.order, w->h_joint_order, w->stream)) ||
      (st = up(J.nc_off, nc_off, w->stream)) || (st = up(J.nc_list, nc_list, w->stream)) ||
      (st = up(J.nc_flag, nc_flag, w->stream)))
    return st;
  B2G_HIP(hipStreamSynchronize(w->stream));
  w->dev_nj = nj;
  w->n_nocollide = int(nc_list.size());
  w->joints_dirty = false;
  return B2G_OK;
}

int upload_world(World* w) {
  if (!w->dirty) return w->joints_dirty ? upload_joints(w) : B2G_OK;
  int st;
  // arrays may be reallocated below (with copies on the default stream): let this world's
  // stream finish first
  B2G_HIP(hipStreamSynchronize(w->stream));
  // contact sets index proxies, which only grow, so they stay valid; their offset arrays
  // are extended with empty segments for the new proxies
  for (int i = 0; i < 2; ++i) {
    ContactArrays& c = w->c[i];
    if (c.off.p && w->set_np[i] > 0 && w->np > w->set_np[i]) {
      if ((st = c.off.reserve(w->np + 1, true, w->set_np[i] + 1))) return st;
    }
  }
  std::vector<float4> force(w->nb, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  if (w->dev_nb > 0) {
    std::vector<float4> f;
    if ((st = down(f, w->b.force, w->dev_nb, w->stream))) return st;
    B2G_HIP(hipStreamSynchronize(w->stream));
    std::copy(f.begin(), f.end(), force.begin());
  }
  if ((st = up(w->b.pos, w->h_pos, w->stream)) || (st = up(w->b.vel, w->h_vel, w->stream)) ||
      (st = up(w->b.pos0, w->h_pos0, w->stream)) || (st = up(w->b.xf, w->h_xf, w->stream)) ||
      (st = up(w->b.mass, w->h_mass, w->stream)) || (st = up(w->b.props, w->h_props, w->stream)) ||
      (st = up(w->b.force, force, w->stream)) || (st = up(w->b.type, w->host_type, w->stream)) ||
      (st = up(w->f.body, w->h_fbody, w->stream)) || (st = up(w->f.shape, w->h_fshape, w->stream)) ||
      (st = up(w->f.mat, w->h_fmat, w->stream)) || (st = up(w->f.filter, w->h_ffilter, w->stream)) ||
      (st = up(w->f.aabb, w->h_faabb, w->stream)) || (st = up(w->f.fat, w->h_ffat, w->stream)) ||
      (st = up(w->f.shapev, w->h_shapev, w->stream)) || (st = up(w->f.world, w->h_fworld, w->stream)) ||
      (st = up(w->b.awake, w->h_awake, w->stream)) || (st = up(w->b.sleep_time, w->h_sleep_time, w->stream)) ||
      (st = up(w->b.auto_sleep, w->h_auto_sleep, w->stream)) || (st = up(w->b.wake, std::vector<int>(w->nb, 0), w->stream)) ||
      (st = up(w->b.bullet, w->h_bullet, w->stream)) ||
      (st = w->f.moved.reserve(std::max(w->np, 1))))
    return st;
  // Big static proxies (ground, container walls: an AABB side over kLooseExtent) stay out of
  // the BVH, where they would inflate every ancestor box; each world's list is tested
  // directly by every emitting proxy of that world. Static fat AABBs never change.
  {
    // (host buffers are world members: the asynchronous copies below read them after return)
    const float kLooseExtent = 8.0f;
    std::vector<int>& tree = w->h_tree;
    std::vector<int>& loose_off = w->h_loose_off;
    std::vector<int>& loose = w->h_loose;
    tree.clear();
    loose.clear();
    loose_off.assign(w->n_worlds + 1, 0);
    std::vector<std::vector<int>> per_world(w->n_worlds);
    for (int f = 0; f < w->np; ++f) {
      float4 a = w->h_ffat[f];
      bool big = host_is_static(w, f) && (a.z - a.x > kLooseExtent || a.w - a.y > kLooseExtent);
      if (big) per_world[w->h_fworld[f]].push_back(f);
      else tree.push_back(f);
    }
    for (int i = 0; i < w->n_worlds; ++i) {
      loose_off[i + 1] = loose_off[i] + int(per_world[i].size());
      loose.insert(loose.end(), per_world[i].begin(), per_world[i].end());
    }
    w->n_tree = int(tree.size());
    if ((st = up(w->f.tree_list, tree, w->stream)) || (st = up(w->f.loose_off, loose_off, w->stream)) ||
        (st = up(w->f.loose, loose, w->stream)))
      return st;
  }
  // per-world solve: the worlds' body ranges (when every world's bodies are contiguous) and the
  // launch order of the worlds, by body-count bucket and larger worlds first within a bucket (the
  // longest workgroups start first)
  {
    const int nwld = w->n_worlds;
    std::vector<int>& off = w->h_world_body_off;
    off.assign(nwld + 1, 0);
    bool contiguous = w->nb > 0;
    for (int b = 1; b < w->nb && contiguous; ++b) contiguous = w->h_bworld[b] >= w->h_bworld[b - 1];
    w->worlds_contiguous = contiguous;
    if (contiguous) {
      for (int b = 0; b < w->nb; ++b) ++off[w->h_bworld[b] + 1];
      for (int i = 0; i < nwld; ++i) off[i + 1] += off[i];
      std::vector<int>& order = w->h_world_order;
      order.resize(nwld);
      for (int i = 0; i < nwld; ++i) order[i] = i;
      auto size = [&](int i) { return off[i + 1] - off[i]; };
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const int ka = world_bucket(size(a)), kb = world_bucket(size(b));
        return ka != kb ? ka < kb : size(a) > size(b);
      });
      for (int k = 0; k < 4; ++k) w->world_bucket_off[k] = 0;
      for (int i = 0; i < nwld; ++i) ++w->world_bucket_off[world_bucket(size(i)) + 1];
      for (int k = 0; k < 3; ++k) w->world_bucket_off[k + 1] += w->world_bucket_off[k];
      if ((st = up(w->s.world_body_off, off, w->stream)) || (st = up(w->s.world_order, order, w->stream))) return st;
    }
  }
  // extend the open contact-set offset arrays: new proxies have empty segments
  for (int i = 0; i < 2; ++i) {
    ContactArrays& c = w->c[i];
    if (c.off.p && w->set_np[i] > 0 && w->np > w->set_np[i]) {
      std::vector<int> last(1);
      B2G_HIP(hipMemcpyAsync(last.data(), c.off.p + w->set_np[i], sizeof(int), hipMemcpyDeviceToHost, w->stream));
      B2G_HIP(hipStreamSynchronize(w->stream));
      std::vector<int> tail(w->np - w->set_np[i], last[0]);
      B2G_HIP(hipMemcpyAsync(c.off.p + w->set_np[i] + 1, tail.data(), sizeof(int) * tail.size(),
                             hipMemcpyHostToDevice, w->stream));
      B2G_HIP(hipStreamSynchronize(w->stream));
      w->set_np[i] = w->np;
    }
  }
  B2G_HIP(hipStreamSynchronize(w->stream));
  w->dev_nb = w->nb;
  w->dev_np = w->np;
  w->dirty = false;
  if (w->nj > 0 || w->joints_dirty) return upload_joints(w);
  return B2G_OK;
}

static int ensure_uploaded(World* w) { return upload_world(w); }

static int run_phase(World* w, int phase, float dt, int vi, int pi) {
  int st;
  if ((st = ensure_uploaded(w))) return st;
  if (phase == B2G_PHASE_COLLIDE) {
    if (w->need_broadphase) {
      if ((st = carry_manifolds(w))) return st;
      if ((st = phase_broadphase(w))) return st;
      w->manifolds_valid[w->cur] = false;
      w->need_broadphase = false;
      // an unfinished sub-stepped TOI pass keeps its cached times on the new set
      if (!w->step_complete && (st = toi_carry_state(w))) return st;
    }
    if ((st = phase_collide(w))) return st;
    w->manifolds_valid[w->cur] = true;
    w->collide_pending = false;
    return B2G_OK;
  }
  if (phase == B2G_PHASE_SOLVE) {
    if (dt <= 0.0f) return B2G_OK;
    float dt_ratio = w->inv_dt0 * dt;
    if ((st = phase_solve(w, dt, dt_ratio, vi, pi))) return st;
    return B2G_OK;
  }
  if (phase == B2G_PHASE_FIND) {
    if (dt > 0.0f) {
      if ((st = phase_synchronize(w))) return st;
      if ((st = carry_manifolds(w))) return st;
      if ((st = phase_broadphase(w))) return st;
      w->manifolds_valid[w->cur] = false;
      w->collide_pending = true;
      w->inv_dt0 = 1.0f / dt;
    }
    return B2G_OK;
  }
  set_error("unknown phase");
  return B2G_ERR_INVALID_ARGUMENT;
}

int kt_harvest(World* w) {
  KernelTimer& t = w->kt;
  const size_t n = t.pending_items.size();
  if (n == 0) return B2G_OK;
  B2G_HIP(hipEventSynchronize(t.ev[2 * n - 1]));
  for (size_t i = 0; i < n; ++i) {
    float ms = 0.0f;
    B2G_HIP(hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]));
    t.ms += ms;
    t.items += t.pending_items[i];
  }
  t.launches += static_cast<long long>(n);
  t.pending_items.clear();
  return B2G_OK;
}

}  // namespace b2g

using namespace b2g;

extern "C" {

const char* b2g_last_error(void) { return g_error.c_str(); }
const char* b2g_version(void) { return "b2g 0.1 (gfx950)"; }

int b2g_world_create(float gx, float gy, b2g_world** out) {
  if (!out) { set_error("null output pointer"); return B2G_ERR_INVALID_ARGUMENT; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("no HIP device available");
    return B2G_ERR_DEVICE;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("hipGetDevice failed");
    return B2G_ERR_DEVICE;
  }
  int st0 = noalloc_scratch_init(dev);  // the no-allocation entry point's scratch on this device
  if (st0) return st0;
  b2g_world* h = new b2g_world();
  World* w = &h->w;
  w->gravity = v2(gx, gy);
  w->device = dev;
  if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess ||
      w->counters.reserve(4 + 128 + 2) != B2G_OK ||
      hipHostMalloc(reinterpret_cast<void**>(&w->h_counters), kPinnedInts * sizeof(int), hipHostMallocDefault) !=
          hipSuccess) {
    set_error("world allocation failed");
    delete h;
    return B2G_ERR_DEVICE;
  }
  for (int i = 0; i < 8; ++i) (void)hipEventCreate(&w->ev[i]);
  std::memset(w->h_counters, 0, kPinnedInts * sizeof(int));
  *out = h;
  return B2G_OK;
}

int b2g_world_destroy(b2g_world* h) {
  if (!h) return B2G_OK;
  World* w = &h->w;
  (void)hipStreamSynchronize(w->stream);
  auto rel = [](auto& a) { a.release(); };
  rel(w->b.pos); rel(w->b.vel); rel(w->b.pos0); rel(w->b.xf); rel(w->b.mass); rel(w->b.props); rel(w->b.force); rel(w->b.awake); rel(w->b.sleep_time); rel(w->b.auto_sleep); rel(w->b.wake);
  rel(w->b.type); rel(w->b.bullet);
  rel(w->f.body); rel(w->f.shape); rel(w->f.mat); rel(w->f.filter); rel(w->f.aabb); rel(w->f.fat); rel(w->f.moved);
  rel(w->f.shapev); rel(w->f.world); rel(w->f.tree_list); rel(w->f.loose_off); rel(w->f.loose);
  for (auto& c : w->c) {
    rel(c.pair); rel(c.off); rel(c.old_index); rel(c.ab); rel(c.hdr); rel(c.geo); rel(c.pt0); rel(c.pt1); rel(c.ids); rel(c.colour);
  }
  for (auto& t : w->toi_state) {
    rel(t.toi); rel(t.count); rel(t.flags);
  }
  rel(w->toi_alpha0); rel(w->toi_ctl); rel(w->toi_big);
  rel(w->toi_cand); rel(w->toi_need); rel(w->toi_rank); rel(w->toi_snap);
  rel(w->toi_cadj_off); rel(w->toi_cadj); rel(w->toi_cadj_cur); rel(w->toi_bfix_off); rel(w->toi_bfix);
  rel(w->toi_bfix_cur); rel(w->toi_cl); rel(w->toi_in_list);
  w->toi_bfix_np = -1;
  rel(w->toi_succ); rel(w->toi_drop_flag); rel(w->toi_drop_pos); rel(w->toi_drop_total);
  {
    ContactArrays& d = w->toi_dropped;
    rel(d.pair); rel(d.off); rel(d.old_index); rel(d.ab); rel(d.hdr); rel(d.geo); rel(d.pt0); rel(d.pt1); rel(d.ids);
    rel(d.colour);
  }
  SolverArrays& s = w->s;
  rel(s.touch_flag); rel(s.touch_scan); rel(s.touch_list); rel(s.k_contact); rel(s.k_bodies); rel(s.k_info);
  rel(s.vnf); rel(s.vm); rel(s.vr0); rel(s.vr1); rel(s.vms); rel(s.vbias); rel(s.K); rel(s.NM); rel(s.lam);
  rel(s.plp); rel(s.pgeo); rel(s.plc); rel(s.prad); rel(s.deg); rel(s.adj_off); rel(s.adj_cursor); rel(s.adj);
  rel(s.color_state); rel(s.color_off); rel(s.color_cursor); rel(s.prev_colour); rel(s.order); rel(s.inv_order); rel(s.parent); rel(s.isl_awake); rel(s.isl_restless);
  rel(s.isl_violated); rel(s.isl_done);
  rel(s.cw_count); rel(s.cw_off); rel(s.cw_cursor); rel(s.world_body_off); rel(s.world_order);
  rel(s.isl_count); rel(s.isl_small); rel(s.isl_start); rel(s.croot);
  BroadphaseArrays& bp = w->bp;
  rel(bp.keys); rel(bp.keys_alt); rel(bp.vals); rel(bp.vals_alt); rel(bp.hist); rel(bp.seg); rel(bp.nodes); rel(bp.nodes4);
  rel(bp.leafrec); rel(bp.count); rel(bp.slots); rel(bp.bounds); rel(bp.scan_tmp); rel(bp.scan_state); rel(bp.scan_ctl); bp.scan_epoch = 0; bp.scan_ctl_ready = false;
  rel(w->counters);
  if (w->h_counters) (void)hipHostFree(w->h_counters);
  for (int i = 0; i < 8; ++i)
    if (w->ev[i]) (void)hipEventDestroy(w->ev[i]);
  for (hipEvent_t e : w->kt.ev) (void)hipEventDestroy(e);
  tape_release(w);
  release_solve_graphs(w);
  if (w->stream) (void)hipStreamDestroy(w->stream);
  delete h;
  return B2G_OK;
}

int b2g_world_set_flags(b2g_world* h, int32_t warm_starting, int32_t allow_sleep, int32_t continuous,
                        int32_t block_solve) {
  if (!h) { set_error("null world"); return B2G_ERR_INVALID_ARGUMENT; }
  if (continuous && h->w.tape.active) {
    set_error("continuous physics cannot be switched on while a tape is recording");
    return B2G_ERR_UNSUPPORTED;
  }
  if (allow_sleep && h->w.tape.active) {
    set_error("allow_sleep cannot be switched on while a tape is recording");
    return B2G_ERR_UNSUPPORTED;
  }
  World* w = &h->w;
  if (w->allow_sleep && !allow_sleep && w->nb > 0) {
    // B2world::set_allow_sleeping(false) wakes every body (src/private/dynamics/b2_world.rs:342-353)
    int st;
    if ((st = ensure_uploaded(w))) return st;
    hipLaunchKernelGGL(k_wake_all, dim3(blocks_for(w->nb, 256)), dim3(256), 0, w->stream, w->nb, w->b.type.p,
                       w->b.awake.p, w->b.sleep_time.p);
    B2G_HIP(hipGetLastError());
    B2G_HIP(hipStreamSynchronize(w->stream));
  }
  w->allow_sleep = allow_sleep != 0;
  w->warm_starting = warm_starting != 0;
  w->continuous = continuous != 0;
  h->w.block_solve = block_solve != 0;
  return B2G_OK;
}

int b2g_world_add(b2g_world* h, int32_t n_bodies, const b2g_body_def* bodies, int32_t n_fixtures,
                  const int32_t* fixture_body, const b2g_shape_def* shapes, const b2g_fixture_def* fixtures,
                  int32_t* proxy_ids) {
  if (!h || n_bodies < 0 || n_fixtures < 0 || (n_bodies > 0 && !bodies) ||
      (n_fixtures > 0 && (!fixture_body || !shapes || !fixtures))) {
    set_error("b2g_world_add: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  int st;
  if (w->dev_nb > 0 && (st = download_state(w))) return st;
  return add_items(w, n_bodies, bodies, nullptr, n_fixtures, fixture_body, shapes, fixtures, proxy_ids);
}

int b2g_world_add_batched(b2g_world* h, int32_t n_bodies, const b2g_body_def* bodies, const int32_t* body_world,
                          int32_t n_fixtures, const int32_t* fixture_body, const b2g_shape_def* shapes,
                          const b2g_fixture_def* fixtures, int32_t* proxy_ids) {
  if (!h || n_bodies < 0 || n_fixtures < 0 || (n_bodies > 0 && (!bodies || !body_world)) ||
      (n_fixtures > 0 && (!fixture_body || !shapes || !fixtures))) {
    set_error("b2g_world_add_batched: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  int st;
  if (w->dev_nb > 0 && (st = download_state(w))) return st;
  return add_items(w, n_bodies, bodies, body_world, n_fixtures, fixture_body, shapes, fixtures, proxy_ids);
}

int b2g_world_world_count(b2g_world* h, int32_t* out) {
  if (!h || !out) { set_error("invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  *out = h->w.n_worlds;
  return B2G_OK;
}

int b2g_world_step(b2g_world* h, float dt, int32_t vi, int32_t pi) {
  if (!h || vi < 0 || pi < 0) { set_error("b2g_world_step: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  int st;
  const bool taping = w->tape.active;
  if (taping) {
    if (dt <= 0.0f || w->dirty || w->nb != w->tape.nb) {
      tape_release(w);
      set_error("b2g_world_step: a recording tape needs dt > 0 and no bodies added while it is open");
      return B2G_ERR_INVALID_ARGUMENT;
    }
    if ((st = tape_step_begin(w))) return st;
  }
  if (w->sub_stepping && w->continuous && w->n_worlds > 1) {
    set_error("b2g_world_step: sub-stepping runs one world, not a batch");
    return B2G_ERR_UNSUPPORTED;
  }
  (void)hipEventRecord(w->ev[0], w->stream);
  if ((st = run_phase(w, B2G_PHASE_COLLIDE, dt, vi, pi))) return st;
  if (taping && (st = tape_after_collide(w))) return st;
  (void)hipEventRecord(w->ev[1], w->stream);
  const float dt_ratio = w->inv_dt0 * dt;
  if (w->step_complete) {
    if ((st = run_phase(w, B2G_PHASE_SOLVE, dt, vi, pi))) return st;
    if (taping && (st = tape_after_solve(w, dt, dt_ratio, vi, pi))) return st;
    (void)hipEventRecord(w->ev[2], w->stream);
    if ((st = run_phase(w, B2G_PHASE_FIND, dt, vi, pi))) return st;
  } else {
    // an unfinished sub-stepped TOI pass: no solve, and so no synchronize or find_new_contacts
    // (src/private/dynamics/b2_world.rs:935-939); the pass resumes below
    (void)hipEventRecord(w->ev[2], w->stream);
    if (dt > 0.0f) w->inv_dt0 = 1.0f / dt;
  }
  w->toi_events = 0;
  if (w->continuous && dt > 0.0f && (st = phase_toi(w, dt, vi))) return st;
  (void)hipEventRecord(w->ev[3], w->stream);
  w->profile_pending = true;
  return B2G_OK;
}

int b2g_world_step_phase(b2g_world* h, int32_t phase, float dt, int32_t vi, int32_t pi) {
  if (!h) { set_error("null world"); return B2G_ERR_INVALID_ARGUMENT; }
  if (h->w.tape.active) {
    set_error("b2g_world_step_phase: close the recording tape first (b2g_tape_end)");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  int st = run_phase(&h->w, phase, dt, vi, pi);
  if (st == B2G_OK && hipStreamSynchronize(h->w.stream) != hipSuccess) {
    set_error("device error");
    return B2G_ERR_DEVICE;
  }
  return st;
}

int b2g_world_body_count(b2g_world* h, int32_t* out) {
  if (!h || !out) { set_error("invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  *out = h->w.nb;
  return B2G_OK;
}
int b2g_world_fixture_count(b2g_world* h, int32_t* out) {
  if (!h || !out) { set_error("invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  *out = h->w.np;
  return B2G_OK;
}
int b2g_world_contact_count(b2g_world* h, int32_t* out) {
  if (!h || !out) { set_error("invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  *out = h->w.n_pairs;
  return B2G_OK;
}

int b2g_world_get_body_states(b2g_world* h, b2g_body_state* out, int32_t n) {
  if (!h || !out || n != h->w.nb) { set_error("b2g_world_get_body_states: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  int st;
  if ((st = ensure_uploaded(w))) return st;
  std::vector<float4> pos, vel, pos0, xf;
  std::vector<int> awake;
  std::vector<float> sleep_time;
  if ((st = down(pos, w->b.pos, n, w->stream)) || (st = down(vel, w->b.vel, n, w->stream)) ||
      (st = down(pos0, w->b.pos0, n, w->stream)) || (st = down(xf, w->b.xf, n, w->stream)) ||
      (st = down(awake, w->b.awake, n, w->stream)) || (st = down(sleep_time, w->b.sleep_time, n, w->stream)))
    return st;
  B2G_HIP(hipStreamSynchronize(w->stream));
  for (int i = 0; i < n; ++i) {
    b2g_body_state& s = out[i];
    s.p[0] = xf[i].x; s.p[1] = xf[i].y; s.q[0] = xf[i].z; s.q[1] = xf[i].w;
    s.c[0] = pos[i].x; s.c[1] = pos[i].y; s.a = pos[i].z;
    s.v[0] = vel[i].x; s.v[1] = vel[i].y; s.w = vel[i].z;
    s.c0[0] = pos0[i].x; s.c0[1] = pos0[i].y; s.a0 = pos0[i].z;
    s.sleep_time = sleep_time[i];
    s.awake = awake[i] ? 1.0f : 0.0f;
    s.pad = 0.0f;
  }
  return B2G_OK;
}

int b2g_world_set_body_states(b2g_world* h, const b2g_body_state* in, int32_t n) {
  if (!h || !in || n != h->w.nb) { set_error("b2g_world_set_body_states: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  int st;
  if ((st = ensure_uploaded(w))) return st;
  std::vector<float4> pos(n), vel(n), pos0(n), xf(n);
  std::vector<int> awake(n);
  std::vector<float> sleep_time(n);
  for (int i = 0; i < n; ++i) {
    const b2g_body_state& s = in[i];
    pos[i] = make_float4(s.c[0], s.c[1], s.a, 0.0f);
    vel[i] = make_float4(s.v[0], s.v[1], s.w, 0.0f);
    pos0[i] = make_float4(s.c0[0], s.c0[1], s.a0, 0.0f);
    xf[i] = make_float4(s.p[0], s.p[1], s.q[0], s.q[1]);
    const bool non_static = w->host_type[i] != B2G_STATIC;
    awake[i] = non_static && s.awake != 0.0f ? 1 : 0;
    sleep_time[i] = non_static ? s.sleep_time : 0.0f;
    if (non_static && !awake[i]) w->created_asleep = true;
  }
  if ((st = up(w->b.pos, pos, w->stream)) || (st = up(w->b.vel, vel, w->stream)) ||
      (st = up(w->b.pos0, pos0, w->stream)) || (st = up(w->b.xf, xf, w->stream)) ||
      (st = up(w->b.awake, awake, w->stream)) || (st = up(w->b.sleep_time, sleep_time, w->stream)))
    return st;
  B2G_HIP(hipStreamSynchronize(w->stream));
  w->h_pos = pos;
  w->h_vel = vel;
  w->h_pos0 = pos0;
  w->h_xf = xf;
  w->h_awake = awake;
  w->h_sleep_time = sleep_time;
  return B2G_OK;
}

int b2g_world_set_fixture_frictions(b2g_world* h, const float* friction, int32_t n) {
  if (!h || !friction || n != h->w.np) {
    set_error("b2g_world_set_fixture_frictions: invalid argument (n must be the fixture count)");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  if (w->tape.active) {
    set_error("b2g_world_set_fixture_frictions: a tape is recording");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  for (int i = 0; i < n; ++i)
    if (!(friction[i] >= 0.0f)) {
      set_error("b2g_world_set_fixture_frictions: frictions must be non-negative");
      return B2G_ERR_INVALID_ARGUMENT;
    }
  int st;
  if ((st = ensure_uploaded(w))) return st;
  for (int i = 0; i < n; ++i) w->h_fmat[i].x = friction[i];
  if ((st = up(w->f.mat, w->h_fmat, w->stream))) return st;
  B2G_HIP(hipStreamSynchronize(w->stream));
  return B2G_OK;
}

int b2g_world_set_bullet(b2g_world* h, int32_t body, int32_t flag) {
  if (!h || body < 0 || body >= h->w.nb) {
    set_error("b2g_world_set_bullet: invalid body");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  w->h_bullet[body] = flag ? 1 : 0;
  w->body_defs[body].bullet = flag ? 1 : 0;
  if (w->b.bullet.p && body < w->dev_nb) {
    B2G_HIP(hipMemcpyAsync(w->b.bullet.p + body, &w->h_bullet[body], sizeof(int), hipMemcpyHostToDevice, w->stream));
    B2G_HIP(hipStreamSynchronize(w->stream));
  }
  return B2G_OK;
}

int b2g_world_set_sub_stepping(b2g_world* h, int32_t flag) {
  if (!h) { set_error("null world"); return B2G_ERR_INVALID_ARGUMENT; }
  h->w.sub_stepping = flag != 0;
  return B2G_OK;
}

int b2g_world_get_sub_stepping(b2g_world* h, int32_t* flag, int32_t* step_complete) {
  if (!h || !flag || !step_complete) { set_error("b2g_world_get_sub_stepping: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  *flag = h->w.sub_stepping ? 1 : 0;
  *step_complete = h->w.step_complete ? 1 : 0;
  return B2G_OK;
}

int b2g_world_get_mass(b2g_world* h, float* out, int32_t n) {
  if (!h || !out || n != h->w.nb) { set_error("b2g_world_get_mass: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  for (int i = 0; i < n; ++i) {
    float* o = out + 6 * i;
    o[0] = w->h_props[i].x;
    o[1] = w->h_mass[i].x;
    o[2] = w->h_I[i];
    o[3] = w->h_mass[i].y;
    o[4] = w->h_mass[i].z;
    o[5] = w->h_mass[i].w;
  }
  return B2G_OK;
}

int b2g_world_get_fat_aabbs(b2g_world* h, float* out, int32_t n) {
  if (!h || !out || n != h->w.np) { set_error("b2g_world_get_fat_aabbs: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  int st;
  if ((st = ensure_uploaded(w))) return st;
  if (n) B2G_HIP(hipMemcpyAsync(out, w->f.fat.p, sizeof(float4) * n, hipMemcpyDeviceToHost, w->stream));
  B2G_HIP(hipStreamSynchronize(w->stream));
  return B2G_OK;
}

int b2g_world_set_fat_aabbs(b2g_world* h, const float* in, int32_t n) {
  if (!h || !in || n != h->w.np) { set_error("b2g_world_set_fat_aabbs: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  int st;
  if ((st = ensure_uploaded(w))) return st;
  if (n) B2G_HIP(hipMemcpyAsync(w->f.fat.p, in, sizeof(float4) * n, hipMemcpyHostToDevice, w->stream));
  B2G_HIP(hipStreamSynchronize(w->stream));
  std::memcpy(w->h_ffat.data(), in, sizeof(float4) * n);
  return B2G_OK;
}

int b2g_world_get_contacts(b2g_world* h, int32_t cap, int32_t* fixtures_ab, int32_t* flags, b2g_manifold* manifolds,
                           int32_t* count) {
  if (!h || !count || cap < 0) { set_error("b2g_world_get_contacts: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  int st;
  if ((st = ensure_uploaded(w))) return st;
  if ((st = carry_manifolds(w))) return st;
  int n = w->n_pairs;
  *count = n;
  int m = std::min(cap, n);
  if (m == 0) return B2G_OK;
  ContactArrays& c = w->c[w->cur];
  std::vector<int2> ab;
  std::vector<int> hdr, tf;
  std::vector<float4> geo, p0, p1;
  std::vector<uint2> ids;
  if ((st = down(ab, c.ab, m, w->stream)) || (st = down(hdr, c.hdr, m, w->stream)) ||
      (st = down(geo, c.geo, m, w->stream)) || (st = down(p0, c.pt0, m, w->stream)) ||
      (st = down(p1, c.pt1, m, w->stream)) || (st = down(ids, c.ids, m, w->stream)))
    return st;
  B2G_HIP(hipStreamSynchronize(w->stream));
  for (int i = 0; i < m; ++i) {
    if (fixtures_ab) { fixtures_ab[2 * i] = ab[i].x; fixtures_ab[2 * i + 1] = ab[i].y; }
    if (flags) flags[i] = (hdr[i] >> 8) > 0 || (hdr[i] & kSensorTouching) ? 1 : 0;
    if (manifolds) {
      b2g_manifold& mf = manifolds[i];
      mf.type = hdr[i] & 0x7f;
      mf.point_count = hdr[i] >> 8;
      mf.local_normal[0] = geo[i].x; mf.local_normal[1] = geo[i].y;
      mf.local_point[0] = geo[i].z; mf.local_point[1] = geo[i].w;
      mf.points[0].local_point[0] = p0[i].x; mf.points[0].local_point[1] = p0[i].y;
      mf.points[0].normal_impulse = p0[i].z; mf.points[0].tangent_impulse = p0[i].w;
      mf.points[0].id = ids[i].x; mf.points[0].pad = 0;
      mf.points[1].local_point[0] = p1[i].x; mf.points[1].local_point[1] = p1[i].y;
      mf.points[1].normal_impulse = p1[i].z; mf.points[1].tangent_impulse = p1[i].w;
      mf.points[1].id = ids[i].y; mf.points[1].pad = 0;
    }
  }
  return B2G_OK;
}

int b2g_world_add_joints(b2g_world* h, int32_t n, const b2g_joint_def* defs) {
  if (!h || n < 0 || (n > 0 && !defs)) { set_error("b2g_world_add_joints: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  if (w->tape.active) { set_error("b2g_world_add_joints: a tape is being recorded"); return B2G_ERR_UNSUPPORTED; }
  const int nj0 = w->nj;
  bool poses = false;
  // the kind of joint k, existing or earlier in this call
  auto kind_of = [&](int k) { return k < nj0 ? w->h_joints[k].type : defs[k - nj0].type; };
  for (int i = 0; i < n; ++i) {
    const b2g_joint_def& d = defs[i];
    const std::string who = "b2g_world_add_joints: joint " + std::to_string(i);
    if (d.type < B2G_JOINT_DISTANCE || d.type > B2G_JOINT_WHEEL) {
      set_error(who + ": unknown kind");
      return B2G_ERR_UNSUPPORTED;
    }
    if (d.body_a < 0 || d.body_a >= w->nb || d.body_b < 0 || d.body_b >= w->nb || d.body_a == d.body_b ||
        w->h_bworld[d.body_a] != w->h_bworld[d.body_b]) {
      set_error(who + ": bad bodies");
      return B2G_ERR_INVALID_ARGUMENT;
    }
    if (d.type == B2G_JOINT_PULLEY && d.ratio == 0.0f) {
      set_error(who + ": a pulley's ratio must not be 0");
      return B2G_ERR_INVALID_ARGUMENT;
    }
    if (d.type == B2G_JOINT_GEAR) {
      // b2_gear_joint::new (src/private/dynamics/joints/b2_gear_joint.rs:8-20): two revolute or
      // prismatic joints created before it
      const int avail = nj0 + i;
      if (d.joint1 < 0 || d.joint1 >= avail || d.joint2 < 0 || d.joint2 >= avail ||
          (kind_of(d.joint1) != B2G_JOINT_REVOLUTE && kind_of(d.joint1) != B2G_JOINT_PRISMATIC) ||
          (kind_of(d.joint2) != B2G_JOINT_REVOLUTE && kind_of(d.joint2) != B2G_JOINT_PRISMATIC)) {
        set_error(who + ": a gear couples two revolute or prismatic joints created before it");
        return B2G_ERR_INVALID_ARGUMENT;
      }
    }
    if (d.type == B2G_JOINT_MOUSE || d.type == B2G_JOINT_GEAR) poses = true;
  }
  int st;
  if (poses && w->dev_nb > 0 && (st = download_state(w))) return st;
  for (int i = 0; i < n; ++i) {
    b2g_joint_def d = defs[i];
    float constant = 0.0f;
    if (d.type == B2G_JOINT_DISTANCE) {
      // b2_distance_joint_new (src/private/dynamics/joints/b2_distance_joint.rs:44-78)
      d.length = std::max(d.length, kLinearSlop);
      d.min_length = std::max(d.min_length, kLinearSlop);
      d.max_length = std::max(d.max_length, d.min_length);
    } else if (d.type == B2G_JOINT_MOUSE) {
      // B2mouseJoint::new (src/joints/b2_mouse_joint.rs:141-150): body B's anchor at the target,
      // b2_mul_t_transform_by_vec2 of body B's transform
      const float4 xf = w->h_xf[d.body_b];
      const float px = d.target[0] - xf.x, py = d.target[1] - xf.y;
      d.local_anchor_b[0] = xf.w * px + xf.z * py;
      d.local_anchor_b[1] = -xf.z * px + xf.w * py;
    } else if (d.type == B2G_JOINT_PULLEY) {
      // b2_pulley_joint::new (src/private/dynamics/joints/b2_pulley_joint.rs:17)
      constant = d.length_a + d.ratio * d.length_b;
    } else if (d.type == B2G_JOINT_GEAR) {
      // the coupled joints' coordinates at the bodies' current poses (b2_gear_joint.rs:21-85)
      float coordinate[2];
      for (int side = 0; side < 2; ++side) {
        const int k = side == 0 ? d.joint1 : d.joint2;
        const b2g_joint_def& c = w->h_joints[k];  // (joints of this call before it are pushed already)
        const float4 pa = w->h_pos[c.body_b], pc = w->h_pos[c.body_a];
        if (c.type == B2G_JOINT_REVOLUTE) {
          coordinate[side] = (pa.z - pc.z) - c.reference_angle;
        } else {
          const float4 xa = w->h_xf[c.body_b], xc = w->h_xf[c.body_a];
          V2 axis = v2(c.local_axis_a[0], c.local_axis_a[1]);
          normalize(&axis);
          const Rot qa{xa.z, xa.w}, qc{xc.z, xc.w};
          const V2 pC = v2(c.local_anchor_a[0], c.local_anchor_a[1]);
          const V2 pA = mulT(qc, mul(qa, v2(c.local_anchor_b[0], c.local_anchor_b[1])) +
                                     (v2(xa.x, xa.y) - v2(xc.x, xc.y)));
          coordinate[side] = dot(pA - pC, axis);
        }
      }
      constant = coordinate[0] + d.ratio * coordinate[1];
    }
    w->h_joints.push_back(d);
    w->h_joint_constant.push_back(constant);
  }
  w->nj = int(w->h_joints.size());
  // colours: first fit over the dynamic bodies (a gear's four), joints in creation order
  std::vector<unsigned long long> used(w->nb, 0ull);
  std::vector<int> colour(w->nj);
  int ncol = 0;
  for (int i = 0; i < w->nj; ++i) {
    const b2g_joint_def& d = w->h_joints[i];
    int jb[4] = {d.body_a, d.body_b, -1, -1};
    if (d.type == B2G_JOINT_GEAR) {
      jb[2] = w->h_joints[d.joint1].body_a;
      jb[3] = w->h_joints[d.joint2].body_a;
    }
    unsigned long long busy = 0ull;
    for (int b : jb)
      if (b >= 0 && w->host_type[b] == B2G_DYNAMIC) busy |= used[b];
    int c = 0;
    while (c < 64 && (busy >> c) & 1ull) ++c;
    if (c == 64) { set_error("b2g_world_add_joints: a body has more than 63 joints"); return B2G_ERR_CAPACITY; }
    colour[i] = c;
    for (int b : jb)
      if (b >= 0 && w->host_type[b] == B2G_DYNAMIC) used[b] |= 1ull << c;
    ncol = std::max(ncol, c + 1);
  }
  w->h_joint_colour_off.assign(ncol + 1, 0);
  for (int i = 0; i < w->nj; ++i) ++w->h_joint_colour_off[colour[i] + 1];
  for (int c = 0; c < ncol; ++c) w->h_joint_colour_off[c + 1] += w->h_joint_colour_off[c];
  std::vector<int> cursor(w->h_joint_colour_off.begin(), w->h_joint_colour_off.end() - 1);
  w->h_joint_order.assign(w->nj, 0);
  for (int i = 0; i < w->nj; ++i) w->h_joint_order[cursor[colour[i]]++] = i;
  w->joints_dirty = true;
  // the next collide works on a contact set found with the joints' collision filter
  // (create_joint flags the bodies' contacts for filtering, src/private/dynamics/b2_world.rs:257-271)
  w->need_broadphase = true;
  return B2G_OK;
}

int b2g_world_set_joint_target(b2g_world* h, int32_t joint, float x, float y) {
  if (!h || joint < 0 || joint >= h->w.nj || h->w.h_joints[joint].type != B2G_JOINT_MOUSE) {
    set_error("b2g_world_set_joint_target: not a mouse joint");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  b2g_joint_def& d = w->h_joints[joint];
  if (d.target[0] == x && d.target[1] == y) return B2G_OK;
  d.target[0] = x;
  d.target[1] = y;
  int st;
  if ((st = ensure_uploaded(w))) return st;
  // the joint row (target in d1.xy) and B2body::set_awake(true) of body B (src/b2_body.rs:783-801)
  const float4 d1 = make_float4(x, y, d.max_force, d.stiffness);
  B2G_HIP(hipMemcpyAsync(w->j.d1.p + joint, &d1, sizeof(float4), hipMemcpyHostToDevice, w->stream));
  const int b = d.body_b;
  if (w->host_type[b] != B2G_STATIC) {
    const int one = 1;
    const float zero = 0.0f;
    B2G_HIP(hipMemcpyAsync(w->b.awake.p + b, &one, sizeof(int), hipMemcpyHostToDevice, w->stream));
    B2G_HIP(hipMemcpyAsync(w->b.sleep_time.p + b, &zero, sizeof(float), hipMemcpyHostToDevice, w->stream));
  }
  B2G_HIP(hipStreamSynchronize(w->stream));
  return B2G_OK;
}

int b2g_world_get_joint_impulses(b2g_world* h, float* out, int32_t n) {
  if (!h || n != h->w.nj || (n > 0 && !out)) {
    set_error("b2g_world_get_joint_impulses: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  int st;
  if ((st = ensure_uploaded(w))) return st;
  std::vector<float4> imp;
  std::vector<float> imp2;
  if ((st = down(imp, w->j.imp, n, w->stream)) || (st = down(imp2, w->j.imp2, n, w->stream))) return st;
  B2G_HIP(hipStreamSynchronize(w->stream));
  for (int i = 0; i < n; ++i) {
    float* o = out + 5 * i;
    o[0] = imp[i].x;
    o[1] = imp[i].y;
    o[2] = imp[i].z;
    o[3] = imp[i].w;
    o[4] = imp2[i];
    // only revolute, prismatic and wheel joints keep limit impulses in the last two slots
    const int t = w->h_joints[i].type;
    if (t != B2G_JOINT_REVOLUTE && t != B2G_JOINT_PRISMATIC && t != B2G_JOINT_WHEEL) o[3] = o[4] = 0.0f;
  }
  return B2G_OK;
}

int b2g_world_set_joint_impulses(b2g_world* h, const float* in, int32_t n) {
  if (!h || n != h->w.nj || (n > 0 && !in)) {
    set_error("b2g_world_set_joint_impulses: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  int st;
  if ((st = ensure_uploaded(w))) return st;
  std::vector<float4> imp(n);
  std::vector<float> imp2(n);
  for (int i = 0; i < n; ++i) {
    const float* o = in + 5 * i;
    imp[i] = make_float4(o[0], o[1], o[2], o[3]);
    imp2[i] = o[4];
  }
  if ((st = up(w->j.imp, imp, w->stream)) || (st = up(w->j.imp2, imp2, w->stream))) return st;
  B2G_HIP(hipStreamSynchronize(w->stream));
  return B2G_OK;
}

int b2g_world_get_shapes(b2g_world* h, b2g_shape_def* out, int32_t n) {
  if (!h || !out || n != h->w.np) { set_error("b2g_world_get_shapes: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  for (int f = 0; f < n; ++f) {
    const HostShape& s = w->h_shapes[f];
    b2g_shape_def& d = out[f];
    std::memset(&d, 0, sizeof(d));
    d.radius = s.radius;
    if (s.type == kCircle) {
      d.type = B2G_SHAPE_CIRCLE;
      d.params[0] = s.p.x;
      d.params[1] = s.p.y;
    } else if (s.type == kEdge) {
      d.type = B2G_SHAPE_EDGE;
      d.flags = s.one_sided ? 1 : 0;
      for (int i = 0; i < 4; ++i) {
        d.params[2 * i] = s.v[i].x;
        d.params[2 * i + 1] = s.v[i].y;
      }
    } else {
      d.type = B2G_SHAPE_POLYGON_EXACT;
      d.count = s.count;
      for (int i = 0; i < s.count; ++i) {
        d.params[2 * i] = s.v[i].x;
        d.params[2 * i + 1] = s.v[i].y;
        d.params[16 + 2 * i] = s.n[i].x;
        d.params[16 + 2 * i + 1] = s.n[i].y;
      }
      d.params[32] = s.p.x;
      d.params[33] = s.p.y;
    }
  }
  return B2G_OK;
}

int b2g_world_get_contact_colours(b2g_world* h, int32_t cap, int32_t* colours, int32_t* count) {
  if (!h || !count || cap < 0 || (cap > 0 && !colours)) {
    set_error("b2g_world_get_contact_colours: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  int st;
  if ((st = ensure_uploaded(w))) return st;
  const int n = w->n_pairs;
  *count = n;
  const int m = std::min(cap, n);
  if (m == 0) return B2G_OK;
  std::fill(colours, colours + m, -1);
  // the previous set's colours, read through each contact's predecessor (as the next step's
  // colouring reads them)
  const int n_old = w->set_np[1 - w->cur] > 0 ? w->n_pairs_old : 0;
  if (n_old == 0 || w->need_broadphase) return B2G_OK;
  std::vector<int> oi, oc;
  if ((st = down(oi, w->c[w->cur].old_index, m, w->stream)) || (st = down(oc, w->c[1 - w->cur].colour, n_old, w->stream)))
    return st;
  B2G_HIP(hipStreamSynchronize(w->stream));
  for (int k = 0; k < m; ++k)
    if (oi[k] >= 0 && oi[k] < n_old) colours[k] = oc[oi[k]];
  return B2G_OK;
}

int b2g_world_get_destroyed_contacts(b2g_world* h, int32_t cap, int32_t* fixtures_ab, b2g_manifold* manifolds,
                                     int32_t* count) {
  if (!h || !count || cap < 0 || (cap > 0 && (!fixtures_ab || !manifolds))) {
    set_error("b2g_world_get_destroyed_contacts: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  int st;
  if ((st = ensure_uploaded(w))) return st;
  *count = 0;
  // before the next collide only: afterwards the predecessor set is that collide's input
  if (w->need_broadphase || !w->collide_pending || w->set_np[1 - w->cur] == 0) return B2G_OK;
  const int n_old = w->n_pairs_old, np = w->n_pairs;
  if (n_old == 0) return B2G_OK;
  const ContactArrays& c = w->c[w->cur];
  const ContactArrays& o = w->c[1 - w->cur];
  std::vector<int> oi, hdr;
  std::vector<int2> ab;
  std::vector<float4> geo, p0, p1;
  std::vector<uint2> ids;
  if ((st = down(oi, c.old_index, np, w->stream)) || (st = down(hdr, o.hdr, n_old, w->stream)) ||
      (st = down(ab, o.ab, n_old, w->stream)) || (st = down(geo, o.geo, n_old, w->stream)) ||
      (st = down(p0, o.pt0, n_old, w->stream)) || (st = down(p1, o.pt1, n_old, w->stream)) ||
      (st = down(ids, o.ids, n_old, w->stream)))
    return st;
  B2G_HIP(hipStreamSynchronize(w->stream));
  std::vector<char> kept(n_old, 0);
  for (int k = 0; k < np; ++k)
    if (oi[k] >= 0 && oi[k] < n_old) kept[oi[k]] = 1;
  int m = 0;
  for (int k = 0; k < n_old; ++k) {
    if (kept[k] || ((hdr[k] >> 8) == 0 && !(hdr[k] & kSensorTouching))) continue;
    if (m < cap) {
      fixtures_ab[2 * m] = ab[k].x;
      fixtures_ab[2 * m + 1] = ab[k].y;
      b2g_manifold& mf = manifolds[m];
      std::memset(&mf, 0, sizeof(mf));
      mf.type = hdr[k] & 0x7f;
      mf.point_count = hdr[k] >> 8;
      mf.local_normal[0] = geo[k].x; mf.local_normal[1] = geo[k].y;
      mf.local_point[0] = geo[k].z; mf.local_point[1] = geo[k].w;
      mf.points[0].local_point[0] = p0[k].x; mf.points[0].local_point[1] = p0[k].y;
      mf.points[0].normal_impulse = p0[k].z; mf.points[0].tangent_impulse = p0[k].w;
      mf.points[0].id = ids[k].x;
      mf.points[1].local_point[0] = p1[k].x; mf.points[1].local_point[1] = p1[k].y;
      mf.points[1].normal_impulse = p1[k].z; mf.points[1].tangent_impulse = p1[k].w;
      mf.points[1].id = ids[k].y;
    }
    ++m;
  }
  *count = m;
  return B2G_OK;
}

int b2g_world_set_contacts(b2g_world* h, int32_t n, const int32_t* fixtures_ab, const b2g_manifold* manifolds,
                           const int32_t* colours) {
  if (!h || n < 0 || (n > 0 && (!fixtures_ab || !manifolds))) {
    set_error("b2g_world_set_contacts: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  int st;
  if ((st = ensure_uploaded(w))) return st;
  if (w->tape.active) { set_error("b2g_world_set_contacts: a tape is being recorded"); return B2G_ERR_INVALID_ARGUMENT; }
  // the pair set of the current fat AABBs becomes the current set ...
  if ((st = phase_broadphase(w))) return st;
  w->need_broadphase = false;
  const int np = w->n_pairs;
  ContactArrays& c = w->c[w->cur];
  ContactArrays& o = w->c[1 - w->cur];
  std::vector<int2> pairs;
  if ((st = down(pairs, c.pair, np, w->stream))) return st;
  B2G_HIP(hipStreamSynchronize(w->stream));
  // ... and the checkpoint becomes its predecessor set: the same pairs, with the checkpoint's
  // manifolds and colours, each contact its own predecessor
  std::unordered_map<long long, int> index;
  index.reserve(size_t(np) * 2 + 1);
  std::vector<int2> ab(np);
  for (int k = 0; k < np; ++k) {
    const int lo = std::min(pairs[k].x, pairs[k].y), hi = std::max(pairs[k].x, pairs[k].y);
    const bool swap = registry_swap(w->h_fshape[lo].x & 0xff, w->h_fshape[hi].x & 0xff) == 1;
    ab[k] = swap ? make_int2(hi, lo) : make_int2(lo, hi);
    index[(long long)ab[k].x << 32 | (unsigned)ab[k].y] = k;
  }
  const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  std::vector<int> hdr(np, 0), col(np, -1), oi(np);
  std::vector<float4> geo(np, z), p0(np, z), p1(np, z);
  std::vector<uint2> ids(np, make_uint2(0u, 0u));
  // entries that are not pairs of the rebuilt set are contacts the next collide destroys: they
  // follow the set's own pairs in the predecessor set (so that their bodies are woken)
  int n_old = np;
  for (int i = 0; i < n; ++i) {
    const int fa = fixtures_ab[2 * i], fb = fixtures_ab[2 * i + 1];
    if (fa < 0 || fb < 0 || fa >= w->np || fb >= w->np || fa == fb) {
      set_error("b2g_world_set_contacts: contact " + std::to_string(i) + " has invalid fixtures");
      return B2G_ERR_INVALID_ARGUMENT;
    }
    const b2g_manifold& m = manifolds[i];
    if (m.point_count < 0 || m.point_count > 2 || m.type < 0 || m.type > 2) {
      set_error("b2g_world_set_contacts: bad manifold " + std::to_string(i));
      return B2G_ERR_INVALID_ARGUMENT;
    }
    auto it = index.find((long long)fa << 32 | (unsigned)fb);
    int k;
    if (it != index.end()) {
      k = it->second;
    } else {
      k = n_old++;
      ab.push_back(make_int2(fa, fb));
      hdr.push_back(0);
      col.push_back(-1);
      geo.push_back(z);
      p0.push_back(z);
      p1.push_back(z);
      ids.push_back(make_uint2(0u, 0u));
      pairs.push_back(make_int2(fa, fb));
    }
    hdr[k] = m.type | (m.point_count << 8);
    geo[k] = make_float4(m.local_normal[0], m.local_normal[1], m.local_point[0], m.local_point[1]);
    p0[k] = make_float4(m.points[0].local_point[0], m.points[0].local_point[1], m.points[0].normal_impulse,
                        m.points[0].tangent_impulse);
    p1[k] = make_float4(m.points[1].local_point[0], m.points[1].local_point[1], m.points[1].normal_impulse,
                        m.points[1].tangent_impulse);
    ids[k] = make_uint2(m.points[0].id, m.points[1].id);
    if (k < np) col[k] = colours ? colours[i] : -1;
  }
  for (int k = 0; k < np; ++k) oi[k] = k;
  const int cap = std::max(n_old, 1);
  if ((st = o.pair.reserve(cap)) || (st = o.ab.reserve(cap)) || (st = o.hdr.reserve(cap)) ||
      (st = o.geo.reserve(cap)) || (st = o.pt0.reserve(cap)) || (st = o.pt1.reserve(cap)) ||
      (st = o.ids.reserve(cap)) || (st = o.colour.reserve(cap)) || (st = o.old_index.reserve(cap)) ||
      (st = o.off.reserve(w->np + 1)) || (st = c.old_index.reserve(std::max(np, 1))))
    return st;
  if (n_old > 0) {
    if ((st = up(o.pair, pairs, w->stream)) || (st = up(o.ab, ab, w->stream)) || (st = up(o.hdr, hdr, w->stream)) ||
        (st = up(o.geo, geo, w->stream)) || (st = up(o.pt0, p0, w->stream)) || (st = up(o.pt1, p1, w->stream)) ||
        (st = up(o.ids, ids, w->stream)) || (st = up(o.colour, col, w->stream)))
      return st;
  }
  if (np > 0 && (st = up(c.old_index, oi, w->stream))) return st;
  B2G_HIP(hipMemcpyAsync(o.off.p, c.off.p, sizeof(int) * (w->np + 1), hipMemcpyDeviceToDevice, w->stream));
  B2G_HIP(hipStreamSynchronize(w->stream));
  w->set_np[1 - w->cur] = w->np;
  w->n_pairs_old = n_old;
  w->manifolds_valid[1 - w->cur] = true;
  w->manifolds_valid[w->cur] = false;
  w->collide_pending = true;
  return B2G_OK;
}

int b2g_world_get_solver_state(b2g_world* h, b2g_solver_state* out) {
  if (!h || !out) { set_error("b2g_world_get_solver_state: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  const World* w = &h->w;
  out->inv_dt0 = w->inv_dt0;
  out->n_colours = w->n_colors;
  out->scratch_colours = w->scratch_colors;
  out->steps_since_scratch = w->steps_since_scratch;
  return B2G_OK;
}

int b2g_world_set_solver_state(b2g_world* h, const b2g_solver_state* in) {
  if (!h || !in || in->n_colours < 0 || in->scratch_colours < 0 || in->steps_since_scratch < 0) {
    set_error("b2g_world_set_solver_state: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  w->inv_dt0 = in->inv_dt0;
  w->n_colors = in->n_colours;
  w->scratch_colors = in->scratch_colours;
  w->steps_since_scratch = in->steps_since_scratch;
  return B2G_OK;
}

int b2g_world_get_solver_order(b2g_world* h, int32_t cap, int32_t* contact_order, int32_t* colour_offsets,
                               int32_t colour_cap, int32_t* n_constraints, int32_t* n_colours) {
  if (!h || !n_constraints || !n_colours) { set_error("invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  *n_constraints = w->n_touch;
  *n_colours = w->n_colors;
  int m = std::min(cap, w->n_touch);
  if (contact_order && m > 0) {
    B2G_HIP(hipMemcpyAsync(contact_order, w->s.k_contact.p, sizeof(int) * m, hipMemcpyDeviceToHost, w->stream));
    B2G_HIP(hipStreamSynchronize(w->stream));
  }
  if (colour_offsets)
    for (int i = 0; i < std::min<int>(colour_cap, int(w->h_color_off.size())); ++i) colour_offsets[i] = w->h_color_off[i];
  return B2G_OK;
}

int b2g_world_get_profile(b2g_world* h, b2g_profile* out) {
  if (!h || !out) { set_error("invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  std::memset(out, 0, sizeof(*out));
  if (w->profile_pending) {
    (void)hipEventSynchronize(w->ev[3]);
    (void)hipEventElapsedTime(&w->prof.collide_ms, w->ev[0], w->ev[1]);
    (void)hipEventElapsedTime(&w->prof.solve_ms, w->ev[1], w->ev[2]);
    (void)hipEventElapsedTime(&w->prof.broadphase_ms, w->ev[2], w->ev[3]);
    (void)hipEventElapsedTime(&w->prof.step_ms, w->ev[0], w->ev[3]);
    w->profile_pending = false;
  }
  out->step = w->prof.step_ms;
  out->collide = w->prof.collide_ms;
  out->solve = w->prof.solve_ms;
  out->broadphase = w->prof.broadphase_ms;
  out->pairs = w->n_pairs;
  out->touching = w->n_touch;
  out->colours = w->n_colors;
  out->position_iterations = w->pos_iters_run;
  out->toi_events = w->toi_events;
  out->island_bins = w->n_bins;
  return B2G_OK;
}

int b2g_world_tune(b2g_world* h, int32_t param, int32_t value) {
  if (!h) { set_error("null world"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  switch (param) {
    case B2G_TUNE_SOLVER_GRAPH:
      w->solver_graphs = value != 0;
      if (!w->solver_graphs) release_solve_graphs(w);
      return B2G_OK;
    case B2G_TUNE_COLOR_ROUNDS:
      if (value < 1 || value > 128) break;
      w->color_rounds = value;
      return B2G_OK;
    case B2G_TUNE_COLOR_REUSE:
      w->color_reuse = value != 0;
      return B2G_OK;
    case B2G_TUNE_WORLD_SOLVE:
      if (value < -1 || value > 1) break;
      w->world_solve = value;
      return B2G_OK;
    case B2G_TUNE_WORLD_THREADS_SMALL:
    case B2G_TUNE_WORLD_THREADS_MEDIUM:
    case B2G_TUNE_WORLD_THREADS_LARGE:
      if (value != 64 && value != 128 && value != 256 && value != 512) break;
      w->world_threads[param - B2G_TUNE_WORLD_THREADS_SMALL] = value;
      return B2G_OK;
    case B2G_TUNE_ISLAND_SOLVE:
      if (value < -1 || value > 1) break;
      w->island_solve = value;
      return B2G_OK;
    case B2G_TUNE_ISLAND_THREADS:
      if (value != 64 && value != 128 && value != 256 && value != 512) break;
      w->bin_threads = value;
      return B2G_OK;
    case B2G_TUNE_TOI_GLOBAL_ISLAND:
      if (value != 0 && value != 1) break;
      w->toi_force_big = value == 1;
      return B2G_OK;
    default: break;
  }
  set_error("b2g_world_tune: unknown parameter or value out of range");
  return B2G_ERR_INVALID_ARGUMENT;
}

int b2g_world_time_kernel(b2g_world* h, int32_t kernel) {
  if (!h || kernel < B2G_KERNEL_NONE || kernel >= B2G_KERNEL_COUNT) {
    set_error("b2g_world_time_kernel: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  int st;
  if ((st = kt_harvest(w))) return st;
  w->kt.kernel = kernel;
  w->kt.ms = 0.0;
  w->kt.launches = w->kt.items = w->kt.aux = 0;
  return B2G_OK;
}

int b2g_world_kernel_time(b2g_world* h, double* ms, int64_t* launches, int64_t* items, int64_t* aux) {
  if (!h || !ms || !launches || !items) {
    set_error("b2g_world_kernel_time: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  int st;
  if ((st = kt_harvest(w))) return st;
  *ms = w->kt.ms;
  *launches = w->kt.launches;
  *items = w->kt.items;
  if (aux) *aux = w->kt.aux;
  w->kt.ms = 0.0;
  w->kt.launches = w->kt.items = w->kt.aux = 0;
  return B2G_OK;
}

int b2g_tape_begin(b2g_world* h, int32_t checkpoint_every) {
  if (!h || checkpoint_every < 1) { set_error("b2g_tape_begin: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  World* w = &h->w;
  if (w->nj > 0) {
    set_error("b2g_tape_begin: the adjoint does not hold joints");
    return B2G_ERR_UNSUPPORTED;
  }
  if (w->continuous || !w->step_complete) {
    // TOI events move bodies outside the taped solve (and an unfinished sub-stepped pass skips it)
    set_error("b2g_tape_begin: the adjoint needs continuous physics off");
    return B2G_ERR_UNSUPPORTED;
  }
  if (w->sleep_possible()) {
    // the frozen model replays every body every step: no sleeping bodies on a tape
    set_error("b2g_tape_begin: the adjoint needs allow_sleep off and every body awake");
    return B2G_ERR_UNSUPPORTED;
  }
  int st;
  if ((st = ensure_uploaded(w))) return st;
  B2G_HIP(hipStreamSynchronize(w->stream));
  tape_release(w);
  w->tape.active = true;
  w->tape.every = checkpoint_every;
  w->tape.nb = w->nb;
  return B2G_OK;
}

int b2g_tape_end(b2g_world* h) {
  if (!h) { set_error("null world"); return B2G_ERR_INVALID_ARGUMENT; }
  h->w.tape.active = false;
  B2G_HIP(hipStreamSynchronize(h->w.stream));
  return B2G_OK;
}

int b2g_tape_free(b2g_world* h) {
  if (!h) { set_error("null world"); return B2G_ERR_INVALID_ARGUMENT; }
  B2G_HIP(hipStreamSynchronize(h->w.stream));
  tape_release(&h->w);
  return B2G_OK;
}

int b2g_tape_info(b2g_world* h, int32_t* steps, int64_t* slots) {
  if (!h || !steps || !slots) { set_error("b2g_tape_info: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  *steps = int32_t(h->w.tape.steps.size());
  *slots = h->w.tape.total;
  return B2G_OK;
}

static int tape_ready(World* w, const char* fn) {
  if (w->tape.active) {
    set_error(std::string(fn) + ": close the recording tape first (b2g_tape_end)");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  if (w->tape.steps.empty()) {
    set_error(std::string(fn) + ": no recorded steps");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  return B2G_OK;
}

int b2g_tape_replay(b2g_world* h, float* final_state) {
  if (!h || !final_state) { set_error("b2g_tape_replay: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  int st;
  if ((st = tape_ready(&h->w, "b2g_tape_replay"))) return st;
  B2G_HIP(hipStreamSynchronize(h->w.stream));
  return tape_replay(&h->w, final_state);
}

int b2g_tape_backward(b2g_world* h, const float* cotangent, float* grad_state, float* grad_friction) {
  if (!h || !cotangent || !grad_state) { set_error("b2g_tape_backward: invalid argument"); return B2G_ERR_INVALID_ARGUMENT; }
  int st;
  if ((st = tape_ready(&h->w, "b2g_tape_backward"))) return st;
  B2G_HIP(hipStreamSynchronize(h->w.stream));
  return tape_backward(&h->w, cotangent, grad_state, grad_friction);
}

int b2g_tape_export(b2g_world* h, int32_t step, int32_t cap, int32_t* slots, float* impulses, int32_t* colour_offsets,
                    int32_t colour_cap, int32_t* n_slots, int32_t* n_colours, float* step_params) {
  if (!h || !n_slots || !n_colours || step < 0 || cap < 0) {
    set_error("b2g_tape_export: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  World* w = &h->w;
  Tape& t = w->tape;
  if (step >= int(t.steps.size())) { set_error("b2g_tape_export: no such step"); return B2G_ERR_INVALID_ARGUMENT; }
  const TapeStep& ts = t.steps[step];
  *n_slots = ts.n;
  *n_colours = int32_t(ts.color_off.size()) - 1;
  if (step_params) {
    step_params[0] = ts.dt; step_params[1] = ts.dt_ratio; step_params[2] = float(ts.vel_iters);
    step_params[3] = float(ts.pos_iters); step_params[4] = ts.warm ? 1.0f : 0.0f; step_params[5] = ts.block ? 1.0f : 0.0f;
  }
  if (colour_offsets)
    for (int i = 0; i < int(ts.color_off.size()) && i < colour_cap; ++i) colour_offsets[i] = ts.color_off[i];
  const int m = std::min(cap, ts.n);
  if (m > 0 && (slots || impulses)) {
    std::vector<int4> sl(m), mt(m);
    std::vector<uint2> id(m);
    std::vector<float4> l0(m), ls(m);
    hipStream_t s = w->stream;
    B2G_HIP(hipMemcpyAsync(sl.data(), t.slot.p + ts.off, sizeof(int4) * m, hipMemcpyDeviceToHost, s));
    B2G_HIP(hipMemcpyAsync(mt.data(), t.meta.p + ts.off, sizeof(int4) * m, hipMemcpyDeviceToHost, s));
    B2G_HIP(hipMemcpyAsync(id.data(), t.ids.p + ts.off, sizeof(uint2) * m, hipMemcpyDeviceToHost, s));
    B2G_HIP(hipMemcpyAsync(l0.data(), t.lam0.p + ts.off, sizeof(float4) * m, hipMemcpyDeviceToHost, s));
    B2G_HIP(hipMemcpyAsync(ls.data(), t.lam_store.p + ts.off, sizeof(float4) * m, hipMemcpyDeviceToHost, s));
    B2G_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < m; ++i) {
      if (slots) {
        int32_t* o = slots + 10 * i;
        o[0] = sl[i].x; o[1] = sl[i].y; o[2] = sl[i].z; o[3] = sl[i].w;
        o[4] = mt[i].x; o[5] = mt[i].y; o[6] = mt[i].z; o[7] = mt[i].w;
        o[8] = int32_t(id[i].x); o[9] = int32_t(id[i].y);
      }
      if (impulses) {
        float* o = impulses + 8 * i;
        o[0] = l0[i].x; o[1] = l0[i].y; o[2] = l0[i].z; o[3] = l0[i].w;
        o[4] = ls[i].x; o[5] = ls[i].y; o[6] = ls[i].z; o[7] = ls[i].w;
      }
    }
  }
  return B2G_OK;
}

int b2g_collide_pairs(b2g_world* h, int32_t n, const int32_t* pairs, const float* xf_a, const float* xf_b,
                      b2g_manifold* out) {
  if (!h || n < 0 || (n > 0 && (!pairs || !xf_a || !xf_b || !out))) {
    set_error("b2g_collide_pairs: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  int st;
  if ((st = ensure_uploaded(&h->w))) return st;
  return collide_fixture_pairs(&h->w, n, pairs, xf_a, xf_b, out);
}

int b2g_collide_pairs_noalloc(b2g_world* h, int32_t n, const int32_t* pairs, const float* xf_a, const float* xf_b,
                              b2g_manifold* out) {
  if (!h || n < 0 || (n > 0 && (!pairs || !xf_a || !xf_b || !out))) {
    set_error("b2g_collide_pairs_noalloc: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  // uploading a world allocates: the world must already be on the device
  if (h->w.dirty || h->w.joints_dirty) {
    set_error("b2g_collide_pairs_noalloc: the world has bodies or joints not yet on the device (step it or call "
              "b2g_collide_pairs once first)");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  return collide_fixture_pairs_noalloc(&h->w, n, pairs, xf_a, xf_b, out);
}

int b2g_broadphase_pairs(b2g_world* h, const float* fat_aabbs, int64_t cap, int32_t* pairs, int64_t* count) {
  if (!h || !fat_aabbs || !count || cap < 0 || (cap > 0 && !pairs)) {
    set_error("b2g_broadphase_pairs: invalid argument");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  int st;
  if ((st = ensure_uploaded(&h->w))) return st;
  return broadphase_pairs_for(&h->w, fat_aabbs, cap, pairs, count);
}

}  // extern "C"
