// b2g adjoint: reverse-mode gradients of a recorded run of steps.
//
// Forward (b2g_world_step with a tape open): after collide, every contact's carried impulses
// and the previous-step contact points they were carried from are noted; after the solve,
// every touching contact becomes a slot (in solve order) holding its bodies, fixtures, manifold
// type/count/feature ids, carry sources, the number of position iterations its island ran and
// the impulses stored in the manifold. Body state (c, a, v, w) is checkpointed every `every`
// steps.
//
// Backward: segments are replayed from their checkpoint with the frozen step (the recorded
// contact set, features, solve order and iteration counts; all continuous quantities
// recomputed by the templated step functions), keeping each step's start state; steps are then
// reversed one by one: the step is replayed once more keeping every constraint's inputs, and
// the vector-Jacobian products of b2g_adjoint_t.h are applied in reverse order -- position
// iterations (colours reversed), position integration, impulse storage, velocity iterations,
// warm start, constraint initialisation and collide geometry, velocity integration. Within a
// colour no two constraints share a dynamic body, so a colour's reverse pass is parallel like
// its forward pass; contributions that several constraints make to one body (through the
// constraint set-up and the manifold geometry) are accumulated with atomics.
#include "b2g_adjoint_t.h"
#include "b2g_internal.h"

namespace b2g {

static constexpr int kBlock = 128;

// ---- recording -------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_tape_pre(int n, int use_old, const int* __restrict__ old_index,
                                                     const int* __restrict__ hdr,
                                                     const uint2* __restrict__ ids, const float4* __restrict__ pt0,
                                                     const float4* __restrict__ pt1, const int* __restrict__ ohdr,
                                                     const uint2* __restrict__ oids, float4* __restrict__ c_lam0,
                                                     int2* __restrict__ c_carry) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int count = hdr[k] >> 8;
  float4 p0 = pt0[k], p1 = pt1[k];
  c_lam0[k] = make_float4(p0.z, p0.w, p1.z, p1.w);
  int carry[2] = {-1, -1};
  const int oi = use_old ? old_index[k] : -1;
  if (oi >= 0 && count > 0) {
    // the same matching as the narrowphase's warm-start copy
    const int ocount = ohdr[oi] >> 8;
    if (ocount > 0) {
      const uint2 oid = oids[oi], id = ids[k];
      for (int i = 0; i < count; ++i) {
        const uint32_t m = i == 0 ? id.x : id.y;
        if (m == oid.x) carry[i] = 2 * oi;
        else if (ocount > 1 && m == oid.y) carry[i] = 2 * oi + 1;
      }
    }
  }
  c_carry[k] = make_int2(carry[0], carry[1]);
}

__global__ void __launch_bounds__(kBlock) k_fill_ll(long long n, long long* __restrict__ a, long long v) {
  long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void __launch_bounds__(kBlock) k_tape_post(int nc, long long off, int has_prev, int pos_iters,
                                                      const int* __restrict__ k_contact, const int2* __restrict__ k_bodies,
                                                      const int4* __restrict__ k_info, const int* __restrict__ isl_done,
                                                      const int2* __restrict__ ab, const int* __restrict__ hdr,
                                                      const uint2* __restrict__ ids, const float4* __restrict__ pt0,
                                                      const float4* __restrict__ pt1, const float4* __restrict__ c_lam0,
                                                      const int2* __restrict__ c_carry,
                                                      const long long* __restrict__ prev_slot,
                                                      long long* __restrict__ cur_slot, int4* __restrict__ t_slot,
                                                      int4* __restrict__ t_meta, uint2* __restrict__ t_ids,
                                                      float4* __restrict__ t_lam0, float4* __restrict__ t_store) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const int k = k_contact[i];
  const long long s = off + i;
  const int2 bd = k_bodies[i], f = ab[k];
  t_slot[s] = make_int4(bd.x, bd.y, f.x, f.y);
  const int2 cc = c_carry[k];
  int src[2] = {-1, -1};
  if (has_prev) {
    if (cc.x >= 0) {
      long long so = prev_slot[cc.x >> 1];
      if (so >= 0) src[0] = static_cast<int>(2 * so + (cc.x & 1));
    }
    if (cc.y >= 0) {
      long long so = prev_slot[cc.y >> 1];
      if (so >= 0) src[1] = static_cast<int>(2 * so + (cc.y & 1));
    }
  }
  const int done = isl_done[k_info[i].w];
  t_meta[s] = make_int4(hdr[k], src[0], src[1], done ? done : pos_iters);
  t_ids[s] = ids[k];
  t_lam0[s] = c_lam0[k];
  const float4 p0 = pt0[k], p1 = pt1[k];
  t_store[s] = make_float4(p0.z, p0.w, p1.z, p1.w);
  cur_slot[k] = s;
}

void tape_release(World* w) {
  Tape& t = w->tape;
  t.slot.release(); t.meta.release(); t.ids.release(); t.lam0.release(); t.lam_store.release();
  t.c_lam0.release(); t.c_carry.release(); t.c_slot[0].release(); t.c_slot[1].release();
  t.ck_pos.release(); t.ck_vel.release();
  t.steps.clear();
  t.total = 0;
  t.ck_count = 0;
  t.active = false;
  t.has_prev = false;
}

int tape_step_begin(World* w) {
  Tape& t = w->tape;
  if (!t.active) return B2G_OK;
  if (t.steps.size() % size_t(t.every) != 0) return B2G_OK;
  const int nb = w->nb;
  int st;
  // a growing array copies its contents outside the world's stream: let the stream finish first
  if (size_t(t.ck_count + 1) * nb > t.ck_pos.cap) B2G_HIP(hipStreamSynchronize(w->stream));
  if ((st = t.ck_pos.reserve(size_t(t.ck_count + 1) * nb, true, size_t(t.ck_count) * nb)) ||
      (st = t.ck_vel.reserve(size_t(t.ck_count + 1) * nb, true, size_t(t.ck_count) * nb)))
    return st;
  B2G_HIP(hipMemcpyAsync(t.ck_pos.p + size_t(t.ck_count) * nb, w->b.pos.p, sizeof(float4) * nb,
                         hipMemcpyDeviceToDevice, w->stream));
  B2G_HIP(hipMemcpyAsync(t.ck_vel.p + size_t(t.ck_count) * nb, w->b.vel.p, sizeof(float4) * nb,
                         hipMemcpyDeviceToDevice, w->stream));
  ++t.ck_count;
  return B2G_OK;
}

int tape_after_collide(World* w) {
  Tape& t = w->tape;
  if (!t.active) return B2G_OK;
  const int n = w->n_pairs;
  int st;
  if (size_t(std::max(n, 1)) > t.c_lam0.cap || size_t(std::max(n, 1)) > t.c_carry.cap)
    B2G_HIP(hipStreamSynchronize(w->stream));
  if ((st = t.c_lam0.reserve(std::max(n, 1))) || (st = t.c_carry.reserve(std::max(n, 1)))) return st;
  if (n == 0) return B2G_OK;
  ContactArrays& c = w->c[w->cur];
  ContactArrays& o = w->c[1 - w->cur];
  // with nothing recorded before, every carried impulse is a constant of the run
  const int use_old = (o.hdr.p && t.has_prev) ? 1 : 0;
  hipLaunchKernelGGL(k_tape_pre, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, w->stream, n, use_old, c.old_index.p,
                     c.hdr.p, c.ids.p, c.pt0.p, c.pt1.p, use_old ? o.hdr.p : c.hdr.p, use_old ? o.ids.p : c.ids.p,
                     t.c_lam0.p, t.c_carry.p);
  B2G_HIP(hipGetLastError());
  return B2G_OK;
}

int tape_after_solve(World* w, float dt, float dt_ratio, int vel_iters, int pos_iters) {
  Tape& t = w->tape;
  if (!t.active) return B2G_OK;
  const int nc = w->n_touch;
  const int np = w->n_pairs;
  if (2 * (t.total + nc) >= (1ll << 31)) {
    set_error("adjoint tape: too many recorded contacts");
    return B2G_ERR_CAPACITY;
  }
  int st;
  const size_t need = size_t(t.total + std::max(nc, 1));
  // a growing array copies its contents outside the world's stream: let the stream finish first
  if (need > t.slot.cap || need > t.meta.cap || need > t.ids.cap || need > t.lam0.cap || need > t.lam_store.cap ||
      size_t(std::max(np, 1)) > t.c_slot[w->cur].cap)
    B2G_HIP(hipStreamSynchronize(w->stream));
  if ((st = t.slot.reserve(need, true, t.total)) || (st = t.meta.reserve(need, true, t.total)) ||
      (st = t.ids.reserve(need, true, t.total)) || (st = t.lam0.reserve(need, true, t.total)) ||
      (st = t.lam_store.reserve(need, true, t.total)) || (st = t.c_slot[w->cur].reserve(std::max(np, 1))) ||
      (st = t.c_slot[1 - w->cur].reserve(1)))
    return st;
  if (np > 0)
    hipLaunchKernelGGL(k_fill_ll, dim3(blocks_for(np, kBlock)), dim3(kBlock), 0, w->stream, static_cast<long long>(np),
                       t.c_slot[w->cur].p, -1ll);
  if (nc > 0) {
    ContactArrays& c = w->c[w->cur];
    SolverArrays& s = w->s;
    hipLaunchKernelGGL(k_tape_post, dim3(blocks_for(nc, kBlock)), dim3(kBlock), 0, w->stream, nc, t.total,
                       int(t.has_prev), pos_iters, s.k_contact.p, s.k_bodies.p, s.k_info.p, s.isl_done.p, c.ab.p,
                       c.hdr.p, c.ids.p, c.pt0.p, c.pt1.p, t.c_lam0.p, t.c_carry.p, t.c_slot[1 - w->cur].p,
                       t.c_slot[w->cur].p, t.slot.p, t.meta.p, t.ids.p, t.lam0.p, t.lam_store.p);
  }
  B2G_HIP(hipGetLastError());
  TapeStep ts;
  ts.n = nc;
  ts.off = t.total;
  if (nc > 0) ts.color_off = w->h_color_off;
  else ts.color_off.assign(1, 0);
  ts.dt = dt;
  ts.dt_ratio = dt_ratio;
  ts.vel_iters = vel_iters;
  ts.pos_iters = pos_iters;
  ts.warm = w->warm_starting;
  ts.block = w->block_solve;
  t.steps.push_back(ts);
  t.total += nc;
  t.has_prev = true;
  return B2G_OK;
}

// ---- replay and reverse ----------------------------------------------------------------------
struct AdjCtx {
  // world
  int nb;
  const int* btype;
  const float4 *bmass, *props, *force;
  float gx, gy;
  const int4* fshape;
  const ShapeV* shapes;
  const float4* fmat;
  // tape
  const int4* slot;
  const int4* meta;
  const uint2* ids;
  const float4* lam0;
  float4* lam_store;
  // step
  long long off;
  int n;
  float dt, dt_ratio;
  int warm, block, record;
  // state
  float4 *pos, *vel;
  const float4 *pos_n, *vel_n;  // at the start of the step
  float4 *vstar, *geom, *lam_cur, *lam_init, *lam0eff;
  float4 *t_ws, *t_vel, *t_ip, *t_pos;
  // cotangents
  float4 *pbar, *vbar, *lambar, *lam0bar, *qnbar, *gbar, *vsbar, *store_bar;
  float *fricbar, *fric_grad;
  int* err;
};

__device__ inline ContactConst contact_const(const AdjCtx& a, int4 sl, int hdr) {
  ContactConst k;
  k.type = hdr & 0xff;
  k.count = hdr >> 8;
  const float4 ma = a.fmat[sl.z], mb = a.fmat[sl.w];
  k.restitution = ma.y > mb.y ? ma.y : mb.y;
  k.threshold = ma.z < mb.z ? ma.z : mb.z;
  k.rA = __int_as_float(a.fshape[sl.z].z);
  k.rB = __int_as_float(a.fshape[sl.w].z);
  const float4 A = a.bmass[sl.x], B = a.bmass[sl.y];
  k.mA = A.x; k.iA = A.y; k.mB = B.x; k.iB = B.y;
  k.lcA[0] = A.z; k.lcA[1] = A.w; k.lcB[0] = B.z; k.lcB[1] = B.w;
  return k;
}

__device__ inline float friction_of(const AdjCtx& a, int4 sl) { return sqrtf(a.fmat[sl.z].x * a.fmat[sl.w].x); }

__device__ inline void put3(float* x, float4 v) { x[0] = v.x; x[1] = v.y; x[2] = v.z; }
__device__ inline float4 f4(const float* x) { return make_float4(x[0], x[1], x[2], 0.0f); }

// the 31 inputs of the velocity-phase functors (b2g_adjoint_t.h)
__device__ inline void vel_inputs(const AdjCtx& a, int i, int4 sl, float4 vA, float4 vB, float4 lam, float* x) {
  put3(x, vA);
  put3(x + 3, vB);
  x[6] = lam.x; x[7] = lam.y; x[8] = lam.z; x[9] = lam.w;
  put3(x + 10, a.pos_n[sl.x]);
  put3(x + 13, a.pos_n[sl.y]);
  const float4 g0 = a.geom[2 * i], g1 = a.geom[2 * i + 1];
  x[16] = g0.x; x[17] = g0.y; x[18] = g0.z; x[19] = g0.w;
  x[20] = g1.x; x[21] = g1.y; x[22] = g1.z; x[23] = g1.w;
  x[24] = friction_of(a, sl);
  put3(x + 25, a.vstar[sl.x]);
  put3(x + 28, a.vstar[sl.y]);
}

__device__ inline TBodyMass<float> body_mass(const AdjCtx& a, int b) {
  const float4 m = a.bmass[b], p = a.props[b];
  TBodyMass<float> r;
  r.inv_m = m.x; r.inv_i = m.y; r.lcx = m.z; r.lcy = m.w;
  r.mass = p.x; r.gravity_scale = p.y; r.lin_damping = p.z; r.ang_damping = p.w;
  return r;
}

// -- forward replay
__global__ void __launch_bounds__(kBlock) k_af_geom(AdjCtx a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const long long s = a.off + i;
  const int4 sl = a.slot[s];
  const int4 mt = a.meta[s];
  const ShapeV &sa = a.shapes[sl.z], &sb = a.shapes[sl.w];
  const float4 A = a.bmass[sl.x], B = a.bmass[sl.y];
  const float lcA[2] = {A.z, A.w}, lcB[2] = {B.z, B.w};
  int hdr = 0;
  uint32_t id[2] = {0, 0};
  GeomFn<float> f{&sa, &sb, lcA, lcB, &hdr, id};
  float x[6], y[8];
  put3(x, a.pos_n[sl.x]);
  put3(x + 3, a.pos_n[sl.y]);
  f(x, y);
  const uint2 rid = a.ids[s];
  if (hdr != mt.x || id[0] != rid.x || id[1] != rid.y) atomicOr(a.err, 1);
  a.geom[2 * i] = make_float4(y[0], y[1], y[2], y[3]);
  a.geom[2 * i + 1] = make_float4(y[4], y[5], y[6], y[7]);
}

__global__ void __launch_bounds__(kBlock) k_af_intvel(AdjCtx a) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nb) return;
  float4 v = a.vel[b];
  if (a.btype[b] == 2) {
    const float4 f = a.force[b];
    IntVelFn<float> fn{body_mass(a, b), f.x, f.y, f.z, a.gx, a.gy, a.dt};
    float x[3], y[3];
    put3(x, v);
    fn(x, y);
    v = f4(y);
    a.vel[b] = v;
  }
  a.vstar[b] = v;
}

__device__ inline float4 carried(const AdjCtx& a, int src, float n0, float t0) {
  if (src < 0) return make_float4(n0, t0, 0.0f, 0.0f);
  const float4 s = a.lam_store[src >> 1];
  return (src & 1) ? make_float4(s.z, s.w, 0.0f, 0.0f) : make_float4(s.x, s.y, 0.0f, 0.0f);
}

__global__ void __launch_bounds__(kBlock) k_af_lam(AdjCtx a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const long long s = a.off + i;
  const int4 mt = a.meta[s];
  const int count = mt.x >> 8;
  const float4 l0 = a.lam0[s];
  const float4 c0 = carried(a, mt.y, l0.x, l0.y), c1 = carried(a, mt.z, l0.z, l0.w);
  const float4 e = make_float4(c0.x, c0.y, c1.x, c1.y);
  a.lam0eff[i] = e;
  float4 li = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (a.warm) {
    li.x = a.dt_ratio * e.x;
    li.y = a.dt_ratio * e.y;
    if (count > 1) {
      li.z = a.dt_ratio * e.z;
      li.w = a.dt_ratio * e.w;
    }
  }
  a.lam_init[i] = li;
  a.lam_cur[i] = li;
}

__global__ void __launch_bounds__(kBlock) k_af_warm(AdjCtx a, int b0, int b1) {
  int i = b0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b1) return;
  const long long s = a.off + i;
  const int4 sl = a.slot[s];
  const ContactConst k = contact_const(a, sl, a.meta[s].x);
  const float4 vA = a.vel[sl.x], vB = a.vel[sl.y];
  if (a.record) {
    a.t_ws[2 * i] = vA;
    a.t_ws[2 * i + 1] = vB;
  }
  float x[kVelIn], y[6];
  vel_inputs(a, i, sl, vA, vB, a.lam_cur[i], x);
  WarmFn<float> fn{k, a.block != 0};
  fn(x, y);
  if (a.btype[sl.x] == 2) a.vel[sl.x] = f4(y);
  if (a.btype[sl.y] == 2) a.vel[sl.y] = f4(y + 3);
}

__global__ void __launch_bounds__(kBlock) k_af_vel(AdjCtx a, int b0, int b1, int it) {
  int i = b0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b1) return;
  const long long s = a.off + i;
  const int4 sl = a.slot[s];
  const ContactConst k = contact_const(a, sl, a.meta[s].x);
  const float4 vA = a.vel[sl.x], vB = a.vel[sl.y], lam = a.lam_cur[i];
  if (a.record) {
    float4* t = a.t_vel + 3 * (size_t(it) * a.n + i);
    t[0] = vA;
    t[1] = vB;
    t[2] = lam;
  }
  float x[kVelIn], y[10];
  vel_inputs(a, i, sl, vA, vB, lam, x);
  VelFn<float> fn{k, a.block != 0};
  fn(x, y);
  if (a.btype[sl.x] == 2) a.vel[sl.x] = f4(y);
  if (a.btype[sl.y] == 2) a.vel[sl.y] = f4(y + 3);
  a.lam_cur[i] = make_float4(y[6], y[7], y[8], y[9]);
}

// velocity point count of a constraint (the block solver may drop to one point)
__device__ inline int vcount_of(const AdjCtx& a, int i, int4 sl, const ContactConst& k) {
  float x[kVelIn];
  vel_inputs(a, i, sl, make_float4(0.0f, 0.0f, 0.0f, 0.0f), make_float4(0.0f, 0.0f, 0.0f, 0.0f),
             make_float4(0.0f, 0.0f, 0.0f, 0.0f), x);
  return velc_of(x, k, a.block != 0).vcount;
}

__global__ void __launch_bounds__(kBlock) k_af_store(AdjCtx a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const long long s = a.off + i;
  const int4 sl = a.slot[s];
  const ContactConst k = contact_const(a, sl, a.meta[s].x);
  const int vc = vcount_of(a, i, sl, k);
  const float4 l = a.lam_cur[i], e = a.lam0eff[i];
  float4 out = make_float4(l.x, l.y, 0.0f, 0.0f);
  if (vc > 1) {
    out.z = l.z;
    out.w = l.w;
  } else if (k.count > 1) {
    out.z = e.z;
    out.w = e.w;
  }
  a.lam_store[s] = out;
}

__global__ void __launch_bounds__(kBlock) k_af_intpos(AdjCtx a) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nb) return;
  if (a.btype[b] == 0) return;
  const float4 p = a.pos[b], v = a.vel[b];
  if (a.record) {
    a.t_ip[2 * b] = p;
    a.t_ip[2 * b + 1] = v;
  }
  IntPosFn<float> fn{a.dt};
  float x[6], y[6];
  put3(x, p);
  put3(x + 3, v);
  fn(x, y);
  a.pos[b] = f4(y);
  a.vel[b] = f4(y + 3);
}

__global__ void __launch_bounds__(kBlock) k_af_pos(AdjCtx a, int b0, int b1, int it) {
  int i = b0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b1) return;
  const long long s = a.off + i;
  const int4 mt = a.meta[s];
  if (mt.w <= it) return;  // the island stopped iterating
  const int4 sl = a.slot[s];
  const ContactConst k = contact_const(a, sl, mt.x);
  const float4 pA = a.pos[sl.x], pB = a.pos[sl.y];
  if (a.record) {
    float4* t = a.t_pos + 2 * (size_t(it) * a.n + i);
    t[0] = pA;
    t[1] = pB;
  }
  float x[14], y[6];
  put3(x, pA);
  put3(x + 3, pB);
  const float4 g0 = a.geom[2 * i], g1 = a.geom[2 * i + 1];
  x[6] = g0.x; x[7] = g0.y; x[8] = g0.z; x[9] = g0.w;
  x[10] = g1.x; x[11] = g1.y; x[12] = g1.z; x[13] = g1.w;
  PosFn<float> fn{k};
  fn(x, y);
  if (a.btype[sl.x] == 2) a.pos[sl.x] = f4(y);
  if (a.btype[sl.y] == 2) a.pos[sl.y] = f4(y + 3);
}

// -- reverse
__device__ inline void add4(float4& d, const float* x) { d.x += x[0]; d.y += x[1]; d.z += x[2]; d.w += x[3]; }
__device__ inline void add3(float4& d, const float* x) { d.x += x[0]; d.y += x[1]; d.z += x[2]; }

__global__ void __launch_bounds__(kBlock) k_ab_clear(AdjCtx a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  a.qnbar[2 * i] = a.qnbar[2 * i + 1] = z;
  a.gbar[2 * i] = a.gbar[2 * i + 1] = z;
  a.vsbar[2 * i] = a.vsbar[2 * i + 1] = z;
  a.fricbar[i] = 0.0f;
}

// The reverse position solve runs kPosGroups threads per constraint, one per group of kPosLanes
// dual lanes over its 14 inputs (adjacent lanes of one wave, as for the velocity solve below:
// all read the position cotangents before any writes, each writes only its own inputs' entries)
constexpr int kPosIn = 14, kPosLanes = 2, kPosGroups = 8;
static_assert(kPosGroups * kPosLanes >= kPosIn && 64 % kPosGroups == 0, "position groups");

__global__ void __launch_bounds__(kBlock) k_ab_pos(AdjCtx a, int b0, int b1, int it) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = b0 + t / kPosGroups, lo = (t % kPosGroups) * kPosLanes;
  if (i >= b1 || lo >= kPosIn) return;
  const long long s = a.off + i;
  const int4 mt = a.meta[s];
  if (mt.w <= it) return;
  const int4 sl = a.slot[s];
  const ContactConst k = contact_const(a, sl, mt.x);
  const bool dA = a.btype[sl.x] == 2, dB = a.btype[sl.y] == 2;
  const float4* tp = a.t_pos + 2 * (size_t(it) * a.n + i);
  float x[kPosIn], yb[6], xb[kPosIn];
  put3(x, tp[0]);
  put3(x + 3, tp[1]);
  const float4 g0 = a.geom[2 * i], g1 = a.geom[2 * i + 1];
  x[6] = g0.x; x[7] = g0.y; x[8] = g0.z; x[9] = g0.w;
  x[10] = g1.x; x[11] = g1.y; x[12] = g1.z; x[13] = g1.w;
  const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  put3(yb, dA ? a.pbar[sl.x] : z);
  put3(yb + 3, dB ? a.pbar[sl.y] : z);
  for (int j = 0; j < kPosIn; ++j) xb[j] = 0.0f;
  vjp_group<kPosIn, 6, kPosLanes>(x, yb, xb, PosFn<float>{k}, lo);
  const int hi = lo + kPosLanes;
  auto put = [&](float* d, int j, bool assign) {
    if (j >= lo && j < hi) {
      if (assign) *d = xb[j];
      else *d += xb[j];
    }
  };
  float* pa = &a.pbar[sl.x].x;
  float* pb = &a.pbar[sl.y].x;
  float* ga = &a.gbar[2 * i].x;
  float* gb = &a.gbar[2 * i + 1].x;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (dA) put(pa + c, c, true);
    if (dB) put(pb + c, 3 + c, true);
  }
  if (lo == 0) {  // the unused fourth component, as a whole-record write would leave it
    if (dA) pa[3] = 0.0f;
    if (dB) pb[3] = 0.0f;
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    put(ga + c, 6 + c, false);
    put(gb + c, 10 + c, false);
  }
}

__global__ void __launch_bounds__(kBlock) k_ab_intpos(AdjCtx a) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nb) return;
  if (a.btype[b] != 2) return;
  float x[6], yb[6], xb[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  put3(x, a.t_ip[2 * b]);
  put3(x + 3, a.t_ip[2 * b + 1]);
  put3(yb, a.pbar[b]);
  put3(yb + 3, a.vbar[b]);
  vjp<6, 6, 6>(x, yb, xb, IntPosFn<float>{a.dt});
  a.pbar[b] = f4(xb);
  a.vbar[b] = f4(xb + 3);
}

__global__ void __launch_bounds__(kBlock) k_ab_store(AdjCtx a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const long long s = a.off + i;
  const int4 sl = a.slot[s];
  const ContactConst k = contact_const(a, sl, a.meta[s].x);
  const int vc = vcount_of(a, i, sl, k);
  const float4 g = a.store_bar[s];
  float4 lb = make_float4(g.x, g.y, 0.0f, 0.0f), eb = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (vc > 1) {
    lb.z = g.z;
    lb.w = g.w;
  } else if (k.count > 1) {
    eb.z = g.z;
    eb.w = g.w;
  }
  a.lambar[i] = lb;
  a.lam0bar[i] = eb;
}

// The reverse velocity solve and warm start run kVelGroups threads per constraint, one per group
// of kVelLanes dual lanes of the vector-Jacobian product (a colour holds a few thousand
// constraints: one thread each would leave most of the GPU idle, and one thread's four passes
// over 8-lane duals run at one wave per SIMD). The groups of a constraint are adjacent lanes of
// one wave: each reads the body and impulse cotangents before any of them writes, and each
// writes only the entries of its own inputs.
constexpr int kVelLanes = 2;
constexpr int kVelGroups = (kVelIn + kVelLanes - 1) / kVelLanes;
static_assert(64 % kVelGroups == 0, "a constraint's groups share one wave");

// cotangents of the velocity inputs of group [lo, lo + kVelLanes): body velocities replaced,
// impulses replaced (kAssignLam) or accumulated, the rest accumulated
template <bool kAssignLam>
__device__ inline void vel_group_out(const AdjCtx& a, int i, int4 sl, bool dA, bool dB, const float* xb, int lo) {
  const int hi = lo + kVelLanes;
  auto put = [&](float* d, int j, bool assign) {
    if (j >= lo && j < hi) {
      if (assign) *d = xb[j];
      else *d += xb[j];
    }
  };
  float* va = &a.vbar[sl.x].x;
  float* vb = &a.vbar[sl.y].x;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (dA) put(va + c, c, true);
    if (dB) put(vb + c, 3 + c, true);
  }
  if (lo == 0) {  // the unused fourth component, as a whole-record write would leave it
    if (dA) va[3] = 0.0f;
    if (dB) vb[3] = 0.0f;
  }
  float* lam = &a.lambar[i].x;
  float* q0 = &a.qnbar[2 * i].x;
  float* q1 = &a.qnbar[2 * i + 1].x;
  float* g0 = &a.gbar[2 * i].x;
  float* g1 = &a.gbar[2 * i + 1].x;
  float* s0 = &a.vsbar[2 * i].x;
  float* s1 = &a.vsbar[2 * i + 1].x;
#pragma unroll
  for (int c = 0; c < 4; ++c) put(lam + c, 6 + c, kAssignLam);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    put(q0 + c, 10 + c, false);
    put(q1 + c, 13 + c, false);
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    put(g0 + c, 16 + c, false);
    put(g1 + c, 20 + c, false);
  }
  put(&a.fricbar[i], 24, false);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    put(s0 + c, 25 + c, false);
    put(s1 + c, 28 + c, false);
  }
}

__global__ void __launch_bounds__(kBlock) k_ab_vel(AdjCtx a, int b0, int b1, int it) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = b0 + t / kVelGroups, lo = (t % kVelGroups) * kVelLanes;
  if (i >= b1) return;
  const long long s = a.off + i;
  const int4 sl = a.slot[s];
  const ContactConst k = contact_const(a, sl, a.meta[s].x);
  const bool dA = a.btype[sl.x] == 2, dB = a.btype[sl.y] == 2;
  const float4* tv = a.t_vel + 3 * (size_t(it) * a.n + i);
  float x[kVelIn], yb[10], xb[kVelIn];
  vel_inputs(a, i, sl, tv[0], tv[1], tv[2], x);
  const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  put3(yb, dA ? a.vbar[sl.x] : z);
  put3(yb + 3, dB ? a.vbar[sl.y] : z);
  const float4 lb = a.lambar[i];
  yb[6] = lb.x; yb[7] = lb.y; yb[8] = lb.z; yb[9] = lb.w;
  for (int j = 0; j < kVelIn; ++j) xb[j] = 0.0f;
  vjp_group<kVelIn, 10, kVelLanes>(x, yb, xb, VelFn<float>{k, a.block != 0}, lo);
  vel_group_out<true>(a, i, sl, dA, dB, xb, lo);
}

__global__ void __launch_bounds__(kBlock) k_ab_warm(AdjCtx a, int b0, int b1) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = b0 + t / kVelGroups, lo = (t % kVelGroups) * kVelLanes;
  if (i >= b1) return;
  const long long s = a.off + i;
  const int4 sl = a.slot[s];
  const ContactConst k = contact_const(a, sl, a.meta[s].x);
  const bool dA = a.btype[sl.x] == 2, dB = a.btype[sl.y] == 2;
  float x[kVelIn], yb[6], xb[kVelIn];
  vel_inputs(a, i, sl, a.t_ws[2 * i], a.t_ws[2 * i + 1], a.lam_init[i], x);
  const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  put3(yb, dA ? a.vbar[sl.x] : z);
  put3(yb + 3, dB ? a.vbar[sl.y] : z);
  for (int j = 0; j < kVelIn; ++j) xb[j] = 0.0f;
  vjp_group<kVelIn, 6, kVelLanes>(x, yb, xb, WarmFn<float>{k, a.block != 0}, lo);
  vel_group_out<false>(a, i, sl, dA, dB, xb, lo);
}

__device__ inline void atomic_add3(float4* p, float4 v) {
  atomicAdd(&p->x, v.x);
  atomicAdd(&p->y, v.y);
  atomicAdd(&p->z, v.z);
}

// carried impulses, friction mixing, manifold geometry, and the constraint set-up's use of
// the start-of-step positions and velocities
__global__ void __launch_bounds__(kBlock) k_ab_setup(AdjCtx a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const long long s = a.off + i;
  const int4 sl = a.slot[s];
  const int4 mt = a.meta[s];
  const int count = mt.x >> 8;
  float4 eb = a.lam0bar[i];
  if (a.warm) {
    const float4 lb = a.lambar[i];
    eb.x += a.dt_ratio * lb.x;
    eb.y += a.dt_ratio * lb.y;
    if (count > 1) {
      eb.z += a.dt_ratio * lb.z;
      eb.w += a.dt_ratio * lb.w;
    }
  }
  if (mt.y >= 0) {
    float* d = &a.store_bar[mt.y >> 1].x + 2 * (mt.y & 1);
    atomicAdd(d, eb.x);
    atomicAdd(d + 1, eb.y);
  }
  if (mt.z >= 0) {
    float* d = &a.store_bar[mt.z >> 1].x + 2 * (mt.z & 1);
    atomicAdd(d, eb.z);
    atomicAdd(d + 1, eb.w);
  }
  // friction = sqrt(fA fB)
  const float fa = a.fmat[sl.z].x, fb = a.fmat[sl.w].x;
  const float mu = sqrtf(fa * fb);
  if (mu > 0.0f) {
    const float g = a.fricbar[i];
    atomicAdd(&a.fric_grad[sl.z], g * 0.5f * fb / mu);
    atomicAdd(&a.fric_grad[sl.w], g * 0.5f * fa / mu);
  }
  // geometry = collide(qA, qB)
  const ShapeV &sa = a.shapes[sl.z], &sb = a.shapes[sl.w];
  const float4 A = a.bmass[sl.x], B = a.bmass[sl.y];
  const float lcA[2] = {A.z, A.w}, lcB[2] = {B.z, B.w};
  float x[6], yb[8], xb[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  put3(x, a.pos_n[sl.x]);
  put3(x + 3, a.pos_n[sl.y]);
  const float4 g0 = a.gbar[2 * i], g1 = a.gbar[2 * i + 1];
  yb[0] = g0.x; yb[1] = g0.y; yb[2] = g0.z; yb[3] = g0.w;
  yb[4] = g1.x; yb[5] = g1.y; yb[6] = g1.z; yb[7] = g1.w;
  vjp<6, 8, 6>(x, yb, xb, GeomFn<float>{&sa, &sb, lcA, lcB, nullptr, nullptr});
  float4 qa = a.qnbar[2 * i], qb = a.qnbar[2 * i + 1];
  qa.x += xb[0]; qa.y += xb[1]; qa.z += xb[2];
  qb.x += xb[3]; qb.y += xb[4]; qb.z += xb[5];
  if (a.btype[sl.x] == 2) {
    atomic_add3(&a.pbar[sl.x], qa);
    atomic_add3(&a.vbar[sl.x], a.vsbar[2 * i]);
  }
  if (a.btype[sl.y] == 2) {
    atomic_add3(&a.pbar[sl.y], qb);
    atomic_add3(&a.vbar[sl.y], a.vsbar[2 * i + 1]);
  }
}

__global__ void __launch_bounds__(kBlock) k_ab_intvel(AdjCtx a) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nb) return;
  if (a.btype[b] != 2) return;
  const float4 f = a.force[b];
  float x[3], yb[3], xb[3] = {0.0f, 0.0f, 0.0f};
  put3(x, a.vel_n[b]);
  put3(yb, a.vbar[b]);
  vjp<3, 3, 3>(x, yb, xb, IntVelFn<float>{body_mass(a, b), f.x, f.y, f.z, a.gx, a.gy, a.dt});
  a.vbar[b] = f4(xb);
}

// ---- drivers ---------------------------------------------------------------------------------
struct AdjBuffers {
  DevArray<float4> pos, vel, seg_pos, seg_vel, vstar, geom, lam_cur, lam_init, lam0eff;
  DevArray<float4> t_ws, t_vel, t_ip, t_pos;
  DevArray<float4> pbar, vbar, lambar, lam0bar, qnbar, gbar, vsbar, store_bar;
  DevArray<float> fricbar, fric_grad;
  DevArray<int> err;
  void release() {
    for (auto* d : {&pos, &vel, &seg_pos, &seg_vel, &vstar, &geom, &lam_cur, &lam_init, &lam0eff, &t_ws, &t_vel, &t_ip,
                    &t_pos, &pbar, &vbar, &lambar, &lam0bar, &qnbar, &gbar, &vsbar, &store_bar})
      d->release();
    fricbar.release();
    fric_grad.release();
    err.release();
  }
};

static AdjCtx make_ctx(World* w, AdjBuffers& B) {
  Tape& t = w->tape;
  AdjCtx a;
  a.nb = w->nb;
  a.btype = w->b.type.p;
  a.bmass = w->b.mass.p;
  a.props = w->b.props.p;
  a.force = w->b.force.p;
  a.gx = w->gravity.x;
  a.gy = w->gravity.y;
  a.fshape = w->f.shape.p;
  a.shapes = w->f.shapev.p;
  a.fmat = w->f.mat.p;
  a.slot = t.slot.p;
  a.meta = t.meta.p;
  a.ids = t.ids.p;
  a.lam0 = t.lam0.p;
  a.lam_store = t.lam_store.p;
  a.off = 0;
  a.n = 0;
  a.dt = a.dt_ratio = 0.0f;
  a.warm = a.block = a.record = 0;
  a.pos = B.pos.p;
  a.vel = B.vel.p;
  a.pos_n = a.vel_n = nullptr;
  a.vstar = B.vstar.p;
  a.geom = B.geom.p;
  a.lam_cur = B.lam_cur.p;
  a.lam_init = B.lam_init.p;
  a.lam0eff = B.lam0eff.p;
  a.t_ws = B.t_ws.p;
  a.t_vel = B.t_vel.p;
  a.t_ip = B.t_ip.p;
  a.t_pos = B.t_pos.p;
  a.pbar = B.pbar.p;
  a.vbar = B.vbar.p;
  a.lambar = B.lambar.p;
  a.lam0bar = B.lam0bar.p;
  a.qnbar = B.qnbar.p;
  a.gbar = B.gbar.p;
  a.vsbar = B.vsbar.p;
  a.store_bar = B.store_bar.p;
  a.fricbar = B.fricbar.p;
  a.fric_grad = B.fric_grad.p;
  a.err = B.err.p;
  return a;
}

static void set_step(AdjCtx& a, const TapeStep& ts, const float4* pos_n, const float4* vel_n, int record) {
  a.off = ts.off;
  a.n = ts.n;
  a.dt = ts.dt;
  a.dt_ratio = ts.dt_ratio;
  a.warm = ts.warm;
  a.block = ts.block;
  a.record = record;
  a.pos_n = pos_n;
  a.vel_n = vel_n;
}

// replays one frozen step on a.pos/a.vel (a.pos_n/a.vel_n: copies of the start state); with
// a.record set, keeps every constraint's inputs for reverse_step
static int replay_step(World* w, const AdjCtx& a, const TapeStep& ts) {
  hipStream_t st = w->stream;
  const int gb = blocks_for(std::max(a.nb, 1), kBlock);
  const int gs = blocks_for(std::max(a.n, 1), kBlock);
  const std::vector<int>& co = ts.color_off;
  const int ncol = int(co.size()) - 1;
  if (a.n > 0) hipLaunchKernelGGL(k_af_geom, dim3(gs), dim3(kBlock), 0, st, a);
  hipLaunchKernelGGL(k_af_intvel, dim3(gb), dim3(kBlock), 0, st, a);
  if (a.n > 0) {
    hipLaunchKernelGGL(k_af_lam, dim3(gs), dim3(kBlock), 0, st, a);
    if (ts.warm)
      for (int c = 0; c < ncol; ++c)
        if (co[c + 1] > co[c])
          hipLaunchKernelGGL(k_af_warm, dim3(blocks_for(co[c + 1] - co[c], kBlock)), dim3(kBlock), 0, st, a, co[c],
                             co[c + 1]);
    for (int it = 0; it < ts.vel_iters; ++it)
      for (int c = 0; c < ncol; ++c)
        if (co[c + 1] > co[c])
          hipLaunchKernelGGL(k_af_vel, dim3(blocks_for(co[c + 1] - co[c], kBlock)), dim3(kBlock), 0, st, a, co[c],
                             co[c + 1], it);
    hipLaunchKernelGGL(k_af_store, dim3(gs), dim3(kBlock), 0, st, a);
  }
  hipLaunchKernelGGL(k_af_intpos, dim3(gb), dim3(kBlock), 0, st, a);
  if (a.n > 0)
    for (int it = 0; it < ts.pos_iters; ++it)
      for (int c = 0; c < ncol; ++c)
        if (co[c + 1] > co[c])
          hipLaunchKernelGGL(k_af_pos, dim3(blocks_for(co[c + 1] - co[c], kBlock)), dim3(kBlock), 0, st, a, co[c],
                             co[c + 1], it);
  B2G_HIP(hipGetLastError());
  return B2G_OK;
}

// reverses one step recorded by replay_step: on entry pbar/vbar hold the cotangents of the
// state after the step and store_bar those of the step's stored impulses; on return pbar/vbar
// hold the cotangents of the start state and store_bar has received those of the previous
// step's stored impulses
static int reverse_step(World* w, const AdjCtx& a, const TapeStep& ts) {
  hipStream_t st = w->stream;
  const int gb = blocks_for(std::max(a.nb, 1), kBlock);
  const int gs = blocks_for(std::max(a.n, 1), kBlock);
  const std::vector<int>& co = ts.color_off;
  const int ncol = int(co.size()) - 1;
  auto range_groups = [&](int c) { return dim3(blocks_for((co[c + 1] - co[c]) * kVelGroups, kBlock)); };
  auto range_pos = [&](int c) { return dim3(blocks_for((co[c + 1] - co[c]) * kPosGroups, kBlock)); };
  if (a.n > 0) {
    hipLaunchKernelGGL(k_ab_clear, dim3(gs), dim3(kBlock), 0, st, a);
    for (int it = ts.pos_iters - 1; it >= 0; --it)
      for (int c = ncol - 1; c >= 0; --c)
        if (co[c + 1] > co[c])
          hipLaunchKernelGGL(k_ab_pos, range_pos(c), dim3(kBlock), 0, st, a, co[c], co[c + 1], it);
  }
  hipLaunchKernelGGL(k_ab_intpos, dim3(gb), dim3(kBlock), 0, st, a);
  if (a.n > 0) {
    hipLaunchKernelGGL(k_ab_store, dim3(gs), dim3(kBlock), 0, st, a);
    for (int it = ts.vel_iters - 1; it >= 0; --it)
      for (int c = ncol - 1; c >= 0; --c)
        if (co[c + 1] > co[c])
          hipLaunchKernelGGL(k_ab_vel, range_groups(c), dim3(kBlock), 0, st, a, co[c], co[c + 1], it);
    if (ts.warm)
      for (int c = ncol - 1; c >= 0; --c)
        if (co[c + 1] > co[c])
          hipLaunchKernelGGL(k_ab_warm, range_groups(c), dim3(kBlock), 0, st, a, co[c], co[c + 1]);
    hipLaunchKernelGGL(k_ab_setup, dim3(gs), dim3(kBlock), 0, st, a);
  }
  hipLaunchKernelGGL(k_ab_intvel, dim3(gb), dim3(kBlock), 0, st, a);
  B2G_HIP(hipGetLastError());
  return B2G_OK;
}

static int reserve_buffers(World* w, AdjBuffers& B, bool backward) {
  Tape& t = w->tape;
  const size_t nb = std::max(w->nb, 1);
  const size_t nf = std::max(w->np, 1);
  size_t maxn = 1;
  int maxv = 1, maxp = 1;
  for (const TapeStep& ts : t.steps) {
    maxn = std::max(maxn, size_t(ts.n));
    maxv = std::max(maxv, ts.vel_iters);
    maxp = std::max(maxp, ts.pos_iters);
  }
  const size_t K = size_t(t.every);
  int st;
  if ((st = B.pos.reserve(nb)) || (st = B.vel.reserve(nb)) || (st = B.vstar.reserve(nb)) ||
      (st = B.geom.reserve(2 * maxn)) || (st = B.lam_cur.reserve(maxn)) || (st = B.lam_init.reserve(maxn)) ||
      (st = B.lam0eff.reserve(maxn)) || (st = B.err.reserve(1)) || (st = B.seg_pos.reserve(K * nb)) ||
      (st = B.seg_vel.reserve(K * nb)))
    return st;
  if (backward) {
    if ((st = B.t_ws.reserve(2 * maxn)) || (st = B.t_vel.reserve(3 * size_t(maxv) * maxn)) ||
        (st = B.t_ip.reserve(2 * nb)) || (st = B.t_pos.reserve(2 * size_t(maxp) * maxn)) ||
        (st = B.pbar.reserve(nb)) || (st = B.vbar.reserve(nb)) || (st = B.lambar.reserve(maxn)) ||
        (st = B.lam0bar.reserve(maxn)) || (st = B.qnbar.reserve(2 * maxn)) || (st = B.gbar.reserve(2 * maxn)) ||
        (st = B.vsbar.reserve(2 * maxn)) || (st = B.store_bar.reserve(size_t(std::max<long long>(t.total, 1)))) ||
        (st = B.fricbar.reserve(maxn)) || (st = B.fric_grad.reserve(nf)))
      return st;
  }
  B2G_HIP(hipMemsetAsync(B.err.p, 0, sizeof(int), w->stream));
  return B2G_OK;
}

static int check_replay(World* w, AdjBuffers& B) {
  int err = 0;
  B2G_HIP(hipMemcpyAsync(&err, B.err.p, sizeof(int), hipMemcpyDeviceToHost, w->stream));
  B2G_HIP(hipStreamSynchronize(w->stream));
  if (err) {
    set_error("adjoint: a recomputed manifold differs from the recorded one (type, point count or features)");
    return B2G_ERR_DEVICE;
  }
  return B2G_OK;
}

static void to_float4(const float* in, std::vector<float4>& p, std::vector<float4>& v, int nb) {
  p.resize(nb);
  v.resize(nb);
  for (int b = 0; b < nb; ++b) {
    const float* x = in + 6 * b;
    p[b] = make_float4(x[0], x[1], x[2], 0.0f);
    v[b] = make_float4(x[3], x[4], x[5], 0.0f);
  }
}

static void from_float4(float* out, const std::vector<float4>& p, const std::vector<float4>& v, int nb) {
  for (int b = 0; b < nb; ++b) {
    float* x = out + 6 * b;
    x[0] = p[b].x; x[1] = p[b].y; x[2] = p[b].z;
    x[3] = v[b].x; x[4] = v[b].y; x[5] = v[b].z;
  }
}

// frozen replay of the whole tape from its first checkpoint; final body state (nb x 6)
int tape_replay(World* w, float* final_state) {
  Tape& t = w->tape;
  if (t.steps.empty()) {
    set_error("adjoint: the tape is empty");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  const int nb = w->nb;
  AdjBuffers B;
  int st = reserve_buffers(w, B, false);
  if (st) { B.release(); return st; }
  AdjCtx a = make_ctx(w, B);
  hipStream_t s = w->stream;
  if (hipMemcpyAsync(B.pos.p, t.ck_pos.p, sizeof(float4) * nb, hipMemcpyDeviceToDevice, s) != hipSuccess ||
      hipMemcpyAsync(B.vel.p, t.ck_vel.p, sizeof(float4) * nb, hipMemcpyDeviceToDevice, s) != hipSuccess) {
    B.release();
    set_error("adjoint: copy failed");
    return B2G_ERR_DEVICE;
  }
  for (const TapeStep& ts : t.steps) {
    if (hipMemcpyAsync(B.seg_pos.p, B.pos.p, sizeof(float4) * nb, hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(B.seg_vel.p, B.vel.p, sizeof(float4) * nb, hipMemcpyDeviceToDevice, s) != hipSuccess) {
      B.release();
      set_error("adjoint: copy failed");
      return B2G_ERR_DEVICE;
    }
    set_step(a, ts, B.seg_pos.p, B.seg_vel.p, 0);
    if ((st = replay_step(w, a, ts))) { B.release(); return st; }
  }
  std::vector<float4> p(nb), v(nb);
  bool ok = hipMemcpyAsync(p.data(), B.pos.p, sizeof(float4) * nb, hipMemcpyDeviceToHost, s) == hipSuccess &&
            hipMemcpyAsync(v.data(), B.vel.p, sizeof(float4) * nb, hipMemcpyDeviceToHost, s) == hipSuccess;
  st = ok ? check_replay(w, B) : B2G_ERR_DEVICE;
  B.release();
  if (!ok) { set_error("adjoint: readback failed"); return B2G_ERR_DEVICE; }
  if (st) return st;
  from_float4(final_state, p, v, nb);
  return B2G_OK;
}

int tape_backward(World* w, const float* cot, float* grad0, float* grad_friction) {
  Tape& t = w->tape;
  if (t.steps.empty()) {
    set_error("adjoint: the tape is empty");
    return B2G_ERR_INVALID_ARGUMENT;
  }
  const int nb = w->nb, nf = w->np;
  const int N = int(t.steps.size());
  const int K = t.every;
  AdjBuffers B;
  int st = reserve_buffers(w, B, true);
  if (st) { B.release(); return st; }
  AdjCtx a = make_ctx(w, B);
  hipStream_t s = w->stream;
  std::vector<float4> hp, hv;
  to_float4(cot, hp, hv, nb);
  auto fail = [&](const char* msg) {
    B.release();
    set_error(msg);
    return B2G_ERR_DEVICE;
  };
  if (hipMemcpyAsync(B.pbar.p, hp.data(), sizeof(float4) * nb, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(B.vbar.p, hv.data(), sizeof(float4) * nb, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(B.store_bar.p, 0, sizeof(float4) * size_t(std::max<long long>(t.total, 1)), s) != hipSuccess ||
      hipMemsetAsync(B.fric_grad.p, 0, sizeof(float) * std::max(nf, 1), s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail("adjoint: upload failed");
  const size_t bytes = sizeof(float4) * nb;
  for (int seg = (N - 1) / K; seg >= 0; --seg) {
    const int n0 = seg * K, n1 = std::min(N, n0 + K);
    if (hipMemcpyAsync(B.pos.p, t.ck_pos.p + size_t(seg) * nb, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(B.vel.p, t.ck_vel.p + size_t(seg) * nb, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return fail("adjoint: copy failed");
    for (int n = n0; n < n1; ++n) {
      float4* sp = B.seg_pos.p + size_t(n - n0) * nb;
      float4* sv = B.seg_vel.p + size_t(n - n0) * nb;
      if (hipMemcpyAsync(sp, B.pos.p, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipMemcpyAsync(sv, B.vel.p, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail("adjoint: copy failed");
      set_step(a, t.steps[n], sp, sv, 0);
      if ((st = replay_step(w, a, t.steps[n]))) { B.release(); return st; }
    }
    for (int n = n1 - 1; n >= n0; --n) {
      float4* sp = B.seg_pos.p + size_t(n - n0) * nb;
      float4* sv = B.seg_vel.p + size_t(n - n0) * nb;
      if (hipMemcpyAsync(B.pos.p, sp, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipMemcpyAsync(B.vel.p, sv, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail("adjoint: copy failed");
      set_step(a, t.steps[n], sp, sv, 1);
      if ((st = replay_step(w, a, t.steps[n]))) { B.release(); return st; }
      if ((st = reverse_step(w, a, t.steps[n]))) { B.release(); return st; }
    }
  }
  std::vector<float> fg(std::max(nf, 1));
  bool ok = hipMemcpyAsync(hp.data(), B.pbar.p, bytes, hipMemcpyDeviceToHost, s) == hipSuccess &&
            hipMemcpyAsync(hv.data(), B.vbar.p, bytes, hipMemcpyDeviceToHost, s) == hipSuccess &&
            hipMemcpyAsync(fg.data(), B.fric_grad.p, sizeof(float) * std::max(nf, 1), hipMemcpyDeviceToHost, s) ==
                hipSuccess;
  st = ok ? check_replay(w, B) : B2G_ERR_DEVICE;
  B.release();
  if (!ok) { set_error("adjoint: readback failed"); return B2G_ERR_DEVICE; }
  if (st) return st;
  from_float4(grad0, hp, hv, nb);
  if (grad_friction) std::copy(fg.begin(), fg.begin() + nf, grad_friction);
  return B2G_OK;
}

}  // namespace b2g
