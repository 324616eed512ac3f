// Continuous collision (time of impact) on the device: the TOI pass that B2world::step runs after
// the solve when continuous physics is on (src/private/dynamics/b2_world.rs:534-901).
//
// The reference loops: find the contact with the smallest time of impact over the contact list,
// advance its two bodies to that time, update the contact, gather the static and kinematic
// contacts of the two bodies into a small island, solve that island's positions (TOI Baumgarte,
// 20 iterations) and velocities over the rest of the step, re-synchronise the bodies' proxies,
// find new pairs, and repeat until no contact has a time of impact before the end of the step.
//
// Here:
// * k_toi_first (grid): at the start of the pass every body's sweep starts at alpha0 = 0, so
//   each contact's time of impact is independent of the others: one thread per contact.
// * k_toi_loop (one persistent workgroup) runs the events (see the device event loop below): the
//   smallest cached time over a candidate list, the event itself sequential as in the reference
//   (advance, update, island gathering over the event bodies' contacts in set order, the island's
//   TOI solve, the new sweeps and velocities), the displaced bodies' TOI flags cleared and proxies
//   re-synchronised, the lost times of impact recomputed. It returns to the host when a proxy was
//   re-fattened: the host rebuilds the contact set (find_new_contacts), lists the contacts whose
//   time is not cached (k_toi_need) and recomputes them (k_toi_rescan: ranked into set order, the
//   reference's sweep alignment on one thread in that order, the times in parallel).
// * Bullets: a contact gets a time of impact when one of its bodies is static, kinematic or a
//   bullet, and a bullet's island gathers dynamic bodies too (src/private/dynamics/b2_world.rs:
//   616-621, 770-777).
// Time of impact: GJK distance (src/private/collision/b2_distance.rs:404-530) and conservative
// advancement with separating axes (src/private/collision/b2_time_of_impact.rs:17-463).
#include "b2g_internal.h"
#include "b2g_collide.h"
#include "b2g_distance.h"
#include "b2g_step_t.h"

namespace b2g {

namespace {

constexpr int kToiBlock = 256;
constexpr int kRescanBlock = 1024;
// A TOI island lives in the event workgroup's LDS when it fits these sizes (bodies: the two plus
// their static/kinematic neighbours; contacts; contacts of one event body looked at when the
// island is gathered), and in a larger store in global memory otherwise.
constexpr int kMaxToiBodies = 64;
constexpr int kMaxToiContacts = 64;
constexpr int kMaxBodyContacts = 512;
constexpr int kBigToiBodies = 4096;
constexpr int kBigToiContacts = 4096;
constexpr int kBigBodyContacts = 16384;
constexpr float kFltEpsilon = 1.1920928955078125e-7f;
constexpr float kSlop = 0.005f;
constexpr float kToiBaumgarte = 0.75f;  // src/b2_common.rs:80
constexpr int kMaxSubSteps = 8;         // src/b2_common.rs:51
constexpr float kTwoPi = 2.0f * 3.14159265359f;

enum : int { kToiValid = 1, kToiDisabled = 2 };
enum : int { kEventNotSolid = 0, kEventSolved = 1, kEventCapacity = 2 };

struct SweepD {
  V2 lc, c0, c;
  float a0, a, alpha0;
};

__device__ inline Xf sweep_xf(const SweepD& s, float beta) {
  Xf xf;
  xf.p = (1.0f - beta) * s.c0 + beta * s.c;
  const float angle = (1.0f - beta) * s.a0 + beta * s.a;
  xf.q = rot(angle);
  xf.p = xf.p - mul(xf.q, s.lc);
  return xf;
}

// B2Sweep::advance (src/b2_math.rs:773-779)
__device__ inline void sweep_advance(SweepD& s, float alpha) {
  const float beta = (alpha - s.alpha0) / (1.0f - s.alpha0);
  s.c0 = s.c0 + beta * (s.c - s.c0);
  s.a0 = s.a0 + beta * (s.a - s.a0);
  s.alpha0 = alpha;
}


struct Separation {
  int type;  // 0 points, 1 face A, 2 face B
  V2 local_point, axis;
};

__device__ Separation separation_init(const Cache& cache, const Proxy& pa, const SweepD& sa, const Proxy& pb,
                                      const SweepD& sb, float t1) {
  Separation f;
  f.local_point = v2(0.0f, 0.0f);
  const Xf xa = sweep_xf(sa, t1), xb = sweep_xf(sb, t1);
  if (cache.count == 1) {
    f.type = 0;
    f.axis = mul(xb, pb.v[cache.ib[0]]) - mul(xa, pa.v[cache.ia[0]]);
    normalize(&f.axis);
  } else if (cache.ia[0] == cache.ia[1]) {
    f.type = 2;
    const V2 b1 = pb.v[cache.ib[0]], b2 = pb.v[cache.ib[1]];
    f.axis = cross(b2 - b1, 1.0f);
    normalize(&f.axis);
    const V2 normal = mul(xb.q, f.axis);
    f.local_point = 0.5f * (b1 + b2);
    if (dot(mul(xa, pa.v[cache.ia[0]]) - mul(xb, f.local_point), normal) < 0.0f) f.axis = -f.axis;
  } else {
    f.type = 1;
    const V2 a1 = pa.v[cache.ia[0]], a2 = pa.v[cache.ia[1]];
    f.axis = cross(a2 - a1, 1.0f);
    normalize(&f.axis);
    const V2 normal = mul(xa.q, f.axis);
    f.local_point = 0.5f * (a1 + a2);
    if (dot(mul(xb, pb.v[cache.ib[0]]) - mul(xa, f.local_point), normal) < 0.0f) f.axis = -f.axis;
  }
  return f;
}

__device__ float separation_min(const Separation& f, int& ia, int& ib, const Proxy& pa, const SweepD& sa,
                                const Proxy& pb, const SweepD& sb, float t) {
  const Xf xa = sweep_xf(sa, t), xb = sweep_xf(sb, t);
  if (f.type == 0) {
    ia = pa.support(mulT(xa.q, f.axis));
    ib = pb.support(mulT(xb.q, -f.axis));
    return dot(mul(xb, pb.v[ib]) - mul(xa, pa.v[ia]), f.axis);
  }
  if (f.type == 1) {
    const V2 normal = mul(xa.q, f.axis);
    const V2 qa = mul(xa, f.local_point);
    ia = -1;
    ib = pb.support(mulT(xb.q, -normal));
    return dot(mul(xb, pb.v[ib]) - qa, normal);
  }
  const V2 normal = mul(xb.q, f.axis);
  const V2 qb = mul(xb, f.local_point);
  ib = -1;
  ia = pa.support(mulT(xa.q, -normal));
  return dot(mul(xa, pa.v[ia]) - qb, normal);
}

__device__ float separation_eval(const Separation& f, int ia, int ib, const Proxy& pa, const SweepD& sa,
                                 const Proxy& pb, const SweepD& sb, float t) {
  const Xf xa = sweep_xf(sa, t), xb = sweep_xf(sb, t);
  if (f.type == 0) return dot(mul(xb, pb.v[ib]) - mul(xa, pa.v[ia]), f.axis);
  if (f.type == 1) return dot(mul(xb, pb.v[ib]) - mul(xa, f.local_point), mul(xa.q, f.axis));
  return dot(mul(xa, pa.v[ia]) - mul(xb, f.local_point), mul(xb.q, f.axis));
}

// b2_time_of_impact with t_max = 1: returns t, *touching set for the touching state
__device__ float time_of_impact(const Proxy& pa, const Proxy& pb, SweepD sa, SweepD sb, bool* touching) {
  *touching = false;
  // B2Sweep::normalize (src/b2_math.rs:782-787)
  float d = kTwoPi * floorf(sa.a0 / kTwoPi);
  sa.a0 -= d;
  sa.a -= d;
  d = kTwoPi * floorf(sb.a0 / kTwoPi);
  sb.a0 -= d;
  sb.a -= d;
  const float target = bmax(kSlop, pa.radius + pb.radius - 3.0f * kSlop);
  const float tolerance = 0.25f * kSlop;
  float t1 = 0.0f;
  Cache cache;
  cache.count = 0;
  cache.metric = 0.0f;
  for (int iter = 0;;) {
    const float dist = gjk_distance(cache, pa, sweep_xf(sa, t1), pb, sweep_xf(sb, t1));
    if (dist <= 0.0f) return 0.0f;  // overlapped
    if (dist < target + tolerance) {
      *touching = true;
      return t1;
    }
    const Separation f = separation_init(cache, pa, sa, pb, sb, t1);
    float t2 = 1.0f;
    for (int push_back = 0;;) {
      int ia = 0, ib = 0;
      float s2 = separation_min(f, ia, ib, pa, sa, pb, sb, t2);
      if (s2 > target + tolerance) return 1.0f;  // separated
      if (s2 > target - tolerance) {
        t1 = t2;
        break;
      }
      float s1 = separation_eval(f, ia, ib, pa, sa, pb, sb, t1);
      if (s1 < target - tolerance) return t1;  // failed
      if (s1 <= target + tolerance) {
        *touching = true;
        return t1;
      }
      float a1 = t1, a2 = t2;
      for (int root = 0;;) {
        float t;
        if (root & 1) t = a1 + (target - s1) * (a2 - a1) / (s2 - s1);
        else t = 0.5f * (a1 + a2);
        ++root;
        const float s = separation_eval(f, ia, ib, pa, sa, pb, sb, t);
        if (fabsf(s - target) < tolerance) {
          t2 = t;
          break;
        }
        if (s > target) {
          a1 = t;
          s1 = s;
        } else {
          a2 = t;
          s2 = s;
        }
        if (root == 50) break;
      }
      if (++push_back == kMaxPolygonVertices) break;
    }
    if (++iter == 20) return t1;  // failed
  }
}

struct ToiArgs {
  int n;                 // contacts in the current set
  int nb;                // bodies
  const int2* ab;        // contact fixtures (registry order)
  int* hdr;              // manifolds of the current set
  float4* geo;
  float4* pt0;
  float4* pt1;
  uint2* ids;
  float* toi;            // per contact: cached alpha, TOI count, flags
  int* tcount;
  int* tflags;
  const int* fbody;      // fixture -> body
  const ShapeV* shapes;  // per fixture
  const float4* fmat;    // friction, restitution, threshold
  const int4* fshape;    // radius bits in .z
  const int* btype;
  const int* bullet;     // per body: B2body's bullet flag
  float4* pos;           // sweep c, a
  float4* pos0;          // sweep c0, a0
  float4* vel;
  float4* xf;
  const float4* bmass;   // inv_mass, inv_I, local centre
  float* alpha0;
  int* awake;            // null: every non-static body is awake
  float* sleep_time;
  unsigned long long* key;  // smallest (alpha bits << 32 | contact)
  int* status;              // event result
  int2* event_bodies;       // the event's two bodies
  int* flags4;              // [0] any proxy re-fattened, [1] need count, [2] / [3] candidate counts
  int* cand;                // contacts of the event's bodies: A's at [0, kMaxBodyContacts), B's after
  int* need;                // contacts whose time of impact is recomputed (cap n)
  int* rank;                // their set-order rank
  float* snap;              // per need entry: both aligned sweeps and alpha0 (19 floats)
  int* moved;               // per proxy: re-fattened by the last synchronisation
  int np;
  int block_solve;
  int velocity_iterations;
  float dt;
};

__device__ inline SweepD load_sweep(const ToiArgs& a, int b) {
  SweepD s;
  const float4 p = a.pos[b], p0 = a.pos0[b], m = a.bmass[b];
  s.lc = v2(m.z, m.w);
  s.c0 = v2(p0.x, p0.y);
  s.a0 = p0.z;
  s.c = v2(p.x, p.y);
  s.a = p.z;
  s.alpha0 = a.alpha0[b];
  return s;
}

__device__ inline void store_sweep(const ToiArgs& a, int b, const SweepD& s) {
  a.pos[b] = make_float4(s.c.x, s.c.y, s.a, 0.0f);
  a.pos0[b] = make_float4(s.c0.x, s.c0.y, s.a0, 0.0f);
  a.alpha0[b] = s.alpha0;
  // synchronize_transform (src/b2_body.rs:974-977)
  const Rot q = rot(s.a);
  const V2 p = s.c - mul(q, s.lc);
  a.xf[b] = make_float4(p.x, p.y, q.s, q.c);
}

__device__ inline bool is_awake(const ToiArgs& a, int b) { return a.btype[b] != 0 && (!a.awake || a.awake[b]); }

__device__ inline void set_awake(const ToiArgs& a, int b) {
  if (a.btype[b] == 0 || !a.awake) return;
  a.awake[b] = 1;
  a.sleep_time[b] = 0.0f;
}

// The reference's conditions for a contact to get a time of impact (src/private/dynamics/
// b2_world.rs:580-621): no sensor, one of its bodies awake, and one static, kinematic or a bullet.
__device__ inline bool toi_candidate(const ToiArgs& a, int k, int* bA, int* bB) {
  const int f = a.tflags[k];
  if ((f & kToiDisabled) || a.tcount[k] > kMaxSubSteps) return false;
  const int2 fab = a.ab[k];
  if (a.fshape[fab.x].w | a.fshape[fab.y].w) return false;
  *bA = a.fbody[fab.x];
  *bB = a.fbody[fab.y];
  if (!is_awake(a, *bA) && !is_awake(a, *bB)) return false;
  return a.btype[*bA] != 2 || a.btype[*bB] != 2 || a.bullet[*bA] || a.bullet[*bB];
}

__device__ inline float toi_alpha(const ToiArgs& a, int k, const SweepD& sa, const SweepD& sb, float alpha0) {
  const int2 fab = a.ab[k];
  bool touching;
  const float beta = time_of_impact(proxy_of(a.shapes[fab.x]), proxy_of(a.shapes[fab.y]), sa, sb, &touching);
  return touching ? bmin(alpha0 + (1.0f - alpha0) * beta, 1.0f) : 1.0f;
}

// First scan of the pass: every alpha0 is 0 (src/private/dynamics/b2_world.rs:546-561).
__global__ void __launch_bounds__(kToiBlock) k_toi_first(ToiArgs a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  for (int b = k; b < a.nb; b += gridDim.x * blockDim.x) a.alpha0[b] = 0.0f;
  if (k >= a.n) return;
  a.tcount[k] = 0;
  a.tflags[k] = 0;
  a.toi[k] = 1.0f;
  int bA, bB;
  if (!toi_candidate(a, k, &bA, &bB)) return;
  SweepD sa = load_sweep(a, bA), sb = load_sweep(a, bB);
  sa.alpha0 = 0.0f;
  sb.alpha0 = 0.0f;
  a.toi[k] = toi_alpha(a, k, sa, sb, 0.0f);
  a.tflags[k] = kToiValid;
}

// A sub-stepped pass resumed by the next step: alpha0, the cached times of impact and the TOI
// counts stay (src/private/dynamics/b2_world.rs:545 resets them only when the last pass was
// complete); the collide in between re-enabled the contacts (src/private/dynamics/b2_contact.rs:
// update), so an event's disabled contact is a candidate again.
__global__ void __launch_bounds__(kToiBlock) k_toi_resume(int n, int* __restrict__ tflags) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) tflags[k] &= kToiValid;
}

// Later scans: contacts whose time of impact is not cached. k_toi_need lists them (any order),
// k_toi_rescan ranks the list into set order, aligns the sweeps on one thread in that order (an
// alignment advances a body's sweep, which later contacts of the same body see) and then computes
// the times of impact in parallel from the aligned sweeps.
// wait_rebuild: list nothing when the last event re-fattened a proxy (the host rebuilds the
// contact set first and lists again)
__global__ void __launch_bounds__(kToiBlock) k_toi_need(ToiArgs a, int wait_rebuild) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (wait_rebuild && a.flags4[0]) return;
  if (k >= a.n || (a.tflags[k] & kToiValid)) return;
  int bA, bB;
  if (toi_candidate(a, k, &bA, &bB)) a.need[atomicAdd(&a.flags4[1], 1)] = k;
}

__device__ void toi_rescan_wg(const ToiArgs& a) {
  const int t = threadIdx.x;
  const int m = a.flags4[1];
  for (int i = t; i < m; i += blockDim.x) {
    const int x = a.need[i];
    int r = 0;
    for (int j = 0; j < m; ++j) r += a.need[j] < x ? 1 : 0;
    a.rank[r] = x;
  }
  __syncthreads();
  if (t == 0) {
    for (int i = 0; i < m; ++i) {
      const int k = a.rank[i];
      const int2 fab = a.ab[k];
      const int ba = a.fbody[fab.x], bb = a.fbody[fab.y];
      SweepD sa = load_sweep(a, ba), sb = load_sweep(a, bb);
      float alpha0 = sa.alpha0;
      if (sa.alpha0 < sb.alpha0) {
        alpha0 = sb.alpha0;
        sweep_advance(sa, alpha0);
        a.pos0[ba] = make_float4(sa.c0.x, sa.c0.y, sa.a0, 0.0f);
        a.alpha0[ba] = sa.alpha0;
      } else if (sb.alpha0 < sa.alpha0) {
        alpha0 = sa.alpha0;
        sweep_advance(sb, alpha0);
        a.pos0[bb] = make_float4(sb.c0.x, sb.c0.y, sb.a0, 0.0f);
        a.alpha0[bb] = sb.alpha0;
      }
      float* q = a.snap + 19 * i;
      const SweepD* sw[2] = {&sa, &sb};
      for (int e = 0; e < 2; ++e) {
        const SweepD& x = *sw[e];
        q[9 * e + 0] = x.lc.x; q[9 * e + 1] = x.lc.y; q[9 * e + 2] = x.c0.x; q[9 * e + 3] = x.c0.y;
        q[9 * e + 4] = x.c.x; q[9 * e + 5] = x.c.y; q[9 * e + 6] = x.a0; q[9 * e + 7] = x.a; q[9 * e + 8] = x.alpha0;
      }
      q[18] = alpha0;
    }
  }
  __syncthreads();
  for (int i = t; i < m; i += blockDim.x) {
    const int k = a.rank[i];
    const float* q = a.snap + 19 * i;
    SweepD sw[2];
    for (int e = 0; e < 2; ++e) {
      sw[e].lc = v2(q[9 * e + 0], q[9 * e + 1]);
      sw[e].c0 = v2(q[9 * e + 2], q[9 * e + 3]);
      sw[e].c = v2(q[9 * e + 4], q[9 * e + 5]);
      sw[e].a0 = q[9 * e + 6];
      sw[e].a = q[9 * e + 7];
      sw[e].alpha0 = q[9 * e + 8];
    }
    a.toi[k] = toi_alpha(a, k, sw[0], sw[1], q[18]);
    a.tflags[k] = kToiValid;
  }
  // ready for the next minimum and the next event's candidate lists
  if (t == 0) {
    *a.key = ~0ull;
    a.flags4[2] = 0;
    a.flags4[3] = 0;
  }
}
__global__ void __launch_bounds__(kRescanBlock) k_toi_rescan(ToiArgs a) { toi_rescan_wg(a); }

// B2contact::update for one contact (src/private/dynamics/b2_contact.rs:103-217) with the bodies'
// current transforms; returns whether it is touching.
__device__ bool update_contact(const ToiArgs& a, int k) {
  const int2 fab = a.ab[k];
  const int bA = a.fbody[fab.x], bB = a.fbody[fab.y];
  const Xf xa = xf_of(a.xf[bA]), xb = xf_of(a.xf[bB]);
  Manifold m;
  evaluate(&m, a.shapes[fab.x], xa, a.shapes[fab.y], xb);
  const int ocount = a.hdr[k] >> 8;
  const uint2 oid = a.ids[k];
  const float4 o0 = a.pt0[k], o1 = a.pt1[k];
  float imp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int i = 0; i < m.count; ++i) {
    if (ocount > 0 && m.id[i] == oid.x) {
      imp[2 * i] = o0.z;
      imp[2 * i + 1] = o0.w;
    } else if (ocount > 1 && m.id[i] == oid.y) {
      imp[2 * i] = o1.z;
      imp[2 * i + 1] = o1.w;
    }
  }
  a.hdr[k] = m.type | (m.count << 8);
  a.geo[k] = make_float4(m.local_normal.x, m.local_normal.y, m.local_point.x, m.local_point.y);
  a.pt0[k] = make_float4(m.lp[0].x, m.lp[0].y, imp[0], imp[1]);
  a.pt1[k] = make_float4(m.lp[1].x, m.lp[1].y, imp[2], imp[3]);
  a.ids[k] = make_uint2(m.id[0], m.id[1]);
  const bool touching = m.count > 0;
  if (touching != (ocount > 0)) {
    set_awake(a, bA);
    set_awake(a, bB);
  }
  a.tflags[k] &= ~kToiDisabled;  // update enables the contact
  return touching;
}

template <int NB, int NC, int NX>
struct EventStore {
  static constexpr int kBodies = NB, kContacts = NC, kCand = NX;
  int bodies[NB];
  SweepD sw[NB];
  int nbody;
  int contacts[NC];
  int ncontact;
  int cand[NX];
  int ncand;
  int overflow;
  TPos<float> P[NB];
  TVel<float> V[NB];
  TVelC<float> vc[NC];
  float lam[NC][4];
  // per island contact, read once: constants for the position solve (TOI masses) and the
  // velocity solve, manifold geometry, friction, island slots
  ContactConst kp[NC], kv[NC];
  TGeom<float> geom[NC];
  float fr[NC];
  int2 slot[NC];
};
using EventShared = EventStore<kMaxToiBodies, kMaxToiContacts, kMaxBodyContacts>;  // LDS
using EventBig = EventStore<kBigToiBodies, kBigToiContacts, kBigBodyContacts>;      // global memory
static_assert(sizeof(EventBig) < (8u << 20), "global island store size");

template <class S>
__device__ inline int island_slot(const S& s, int b) {
  for (int i = 0; i < s.nbody; ++i)
    if (s.bodies[i] == b) return i;
  return -1;
}

template <class S>
__device__ inline ContactConst contact_const(const ToiArgs& a, int k, const S& s, bool toi_masses,
                                             int* ia, int* ib, float* friction, TGeom<float>* g) {
  const int2 fab = a.ab[k];
  const int bA = a.fbody[fab.x], bB = a.fbody[fab.y];
  *ia = island_slot(s, bA);
  *ib = island_slot(s, bB);
  const float4 matA = a.fmat[fab.x], matB = a.fmat[fab.y];
  ContactConst c;
  const int h = a.hdr[k];
  c.type = h & 0xff;
  c.count = h >> 8;
  // mixing laws, src/b2_contact.rs:29-46
  *friction = sqrtf(matA.x * matB.x);
  c.restitution = matA.y > matB.y ? matA.y : matB.y;
  c.threshold = matA.z < matB.z ? matA.z : matB.z;
  c.rA = __int_as_float(a.fshape[fab.x].z);
  c.rB = __int_as_float(a.fshape[fab.y].z);
  const float4 mA = a.bmass[bA], mB = a.bmass[bB];
  c.mA = mA.x;
  c.iA = mA.y;
  c.mB = mB.x;
  c.iB = mB.y;
  // the TOI position solve moves only the two TOI bodies (island slots 0 and 1)
  if (toi_masses && *ia > 1) c.mA = c.iA = 0.0f;
  if (toi_masses && *ib > 1) c.mB = c.iB = 0.0f;
  c.lcA[0] = mA.z;
  c.lcA[1] = mA.w;
  c.lcB[0] = mB.z;
  c.lcB[1] = mB.w;
  const float4 gg = a.geo[k], p0 = a.pt0[k], p1 = a.pt1[k];
  g->ln = v2(gg.x, gg.y);
  g->lpt = v2(gg.z, gg.w);
  g->lp[0] = v2(p0.x, p0.y);
  g->lp[1] = v2(p1.x, p1.y);
  return c;
}

struct EventCtl {
  float min_alpha;
  int min_k, solid;
};

template <class S>
__device__ void toi_event_island(const ToiArgs& a, S& s, float min_alpha);

// One TOI event (src/private/dynamics/b2_world.rs:683-898), called by a whole workgroup (the
// event itself runs on its first thread, as sequential as the reference's). S: the island's
// store (LDS or global memory).
template <class S>
__device__ void toi_event(const ToiArgs& a, S& s, EventCtl& e) {
  const int t = threadIdx.x;
  if (t == 0) {
    float& min_alpha = e.min_alpha;
    int& min_k = e.min_k;
    int& solid = e.solid;
    const unsigned long long key = *a.key;
    min_k = static_cast<int>(key & 0xffffffffu);
    min_alpha = __uint_as_float(static_cast<unsigned>(key >> 32));
    s.overflow = 0;
    const int k = min_k;
    const int2 fab = a.ab[k];
    const int bA = a.fbody[fab.x], bB = a.fbody[fab.y];
    const SweepD backup1 = load_sweep(a, bA), backup2 = load_sweep(a, bB);
    // B2body::advance (src/b2_body.rs:979-986)
    SweepD sa = backup1, sb = backup2;
    sweep_advance(sa, min_alpha);
    sa.c = sa.c0;
    sa.a = sa.a0;
    store_sweep(a, bA, sa);
    sweep_advance(sb, min_alpha);
    sb.c = sb.c0;
    sb.a = sb.a0;
    store_sweep(a, bB, sb);
    *a.event_bodies = make_int2(bA, bB);
    const bool touching = update_contact(a, k);
    a.tflags[k] &= ~kToiValid;
    a.tcount[k] += 1;
    solid = touching ? 1 : 0;
    if (!touching) {
      a.tflags[k] |= kToiDisabled;
      store_sweep(a, bA, backup1);
      store_sweep(a, bB, backup2);
      *a.status = kEventNotSolid;
    } else {
      set_awake(a, bA);
      set_awake(a, bB);
      s.bodies[0] = bA;
      s.sw[0] = sa;
      s.bodies[1] = bB;
      s.sw[1] = sb;
      s.nbody = 2;
      s.contacts[0] = k;
      s.ncontact = 1;
    }
  }
  __syncthreads();
  if (e.solid && t == 0) toi_event_island(a, s, e.min_alpha);
  __syncthreads();
}

// the solid event's island: its static and kinematic contacts gathered, its TOI solve, the new
// sweeps and velocities (first thread only)
template <class S>
__device__ void toi_event_island(const ToiArgs& a, S& s, float min_alpha) {
  // gather the static and kinematic contacts of each dynamic TOI body, in set order
  for (int side = 0; side < 2; ++side) {
    const int body = s.bodies[side];
    if (a.btype[body] != 2) continue;
    const int nc = a.flags4[2 + side];
    if (nc > S::kCand) {
      s.overflow = 1;
      break;
    }
    s.ncand = nc;
    for (int i = 0; i < nc; ++i) s.cand[i] = a.cand[side * S::kCand + i];
    for (int i = 1; i < nc; ++i) {
      const int x = s.cand[i];
      int j = i - 1;
      while (j >= 0 && s.cand[j] > x) {
        s.cand[j + 1] = s.cand[j];
        --j;
      }
      s.cand[j + 1] = x;
    }
    for (int i = 0; i < s.ncand; ++i) {
      const int k = s.cand[i];
      bool in = false;
      for (int j = 0; j < s.ncontact; ++j) in = in || s.contacts[j] == k;
      if (in) continue;
      const int2 fab = a.ab[k];
      const int other = a.fbody[fab.x] == body ? a.fbody[fab.y] : a.fbody[fab.x];
      // only static, kinematic or bullet bodies join (src/private/dynamics/b2_world.rs:770-777)
      if (a.btype[other] == 2 && !a.bullet[body] && !a.bullet[other]) continue;
      if (a.fshape[fab.x].w | a.fshape[fab.y].w) continue;  // sensors (784-789)
      const int slot = island_slot(s, other);
      const SweepD backup = load_sweep(a, other);
      SweepD so = backup;
      if (slot < 0) {
        sweep_advance(so, min_alpha);
        so.c = so.c0;
        so.a = so.a0;
        store_sweep(a, other, so);
      }
      if (!update_contact(a, k)) {
        store_sweep(a, other, backup);
        continue;
      }
      if (s.ncontact == S::kContacts) {
        s.overflow = 1;
        break;
      }
      s.contacts[s.ncontact++] = k;
      if (slot >= 0) continue;
      if (s.nbody == S::kBodies) {
        s.overflow = 1;
        break;
      }
      set_awake(a, other);
      s.bodies[s.nbody] = other;
      s.sw[s.nbody] = so;
      ++s.nbody;
    }
    if (s.overflow) break;
  }
  if (s.overflow) {
    *a.status = kEventCapacity;
    return;
  }
  // B2island::solve_toi (src/private/dynamics/b2_island_private.rs:330-458)
  const float h = (1.0f - min_alpha) * a.dt;
  for (int i = 0; i < s.nbody; ++i) {
    const int b = s.bodies[i];
    s.P[i].c = s.sw[i].c;
    s.P[i].a = s.sw[i].a;
    const float4 v = a.vel[b];
    s.V[i].v = v2(v.x, v.y);
    s.V[i].w = v.z;
  }
  for (int j = 0; j < s.ncontact; ++j) {
    int ia, ib;
    s.kv[j] = contact_const(a, s.contacts[j], s, false, &ia, &ib, &s.fr[j], &s.geom[j]);
    s.kp[j] = s.kv[j];
    // the TOI position solve moves only the two TOI bodies (island slots 0 and 1)
    if (ia > 1) s.kp[j].mA = s.kp[j].iA = 0.0f;
    if (ib > 1) s.kp[j].mB = s.kp[j].iB = 0.0f;
    s.slot[j] = make_int2(ia, ib);
  }
  for (int it = 0; it < 20; ++it) {
    float min_sep = 0.0f;
    for (int j = 0; j < s.ncontact; ++j)
      min_sep = bmin(min_sep, solve_position_t(s.kp[j], s.geom[j], s.P[s.slot[j].x], s.P[s.slot[j].y],
                                               double(kToiBaumgarte)));
    if (min_sep >= -1.5f * kSlop) break;
  }
  s.sw[0].c0 = s.P[0].c;
  s.sw[0].a0 = s.P[0].a;
  s.sw[1].c0 = s.P[1].c;
  s.sw[1].a0 = s.P[1].a;
  for (int j = 0; j < s.ncontact; ++j) {
    const int2 sl = s.slot[j];
    s.vc[j] = init_velocity_t(s.kv[j], s.geom[j], s.fr[j], s.P[sl.x], s.P[sl.y], s.V[sl.x], s.V[sl.y],
                              a.block_solve != 0);
    s.lam[j][0] = s.lam[j][1] = s.lam[j][2] = s.lam[j][3] = 0.0f;
  }
  for (int it = 0; it < a.velocity_iterations; ++it)
    for (int j = 0; j < s.ncontact; ++j)
      solve_velocity_t(s.vc[j], s.kv[j], s.lam[j], s.V[s.slot[j].x], s.V[s.slot[j].y], a.block_solve != 0);
  for (int i = 0; i < s.nbody; ++i) {
    integrate_position_t(s.P[i], s.V[i], h);
    const int b = s.bodies[i];
    SweepD sw = s.sw[i];
    sw.c = s.P[i].c;
    sw.a = s.P[i].a;
    store_sweep(a, b, sw);
    a.vel[b] = make_float4(s.V[i].v.x, s.V[i].v.y, s.V[i].w, 0.0f);
  }
  *a.status = kEventSolved;
}

// TOI state of a rebuilt contact set from its predecessors (new pairs start uncached).
__global__ void __launch_bounds__(kToiBlock) k_toi_carry(int n, const int* __restrict__ old_index,
                                                         const float* __restrict__ otoi, const int* __restrict__ ocount,
                                                         const int* __restrict__ oflags, float* __restrict__ toi,
                                                         int* __restrict__ tcount, int* __restrict__ tflags) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int o = old_index[k];
  toi[k] = o >= 0 ? otoi[o] : 1.0f;
  tcount[k] = o >= 0 ? ocount[o] : 0;
  tflags[k] = o >= 0 ? oflags[o] : 0;
}

// ---- device event loop ---------------------------------------------------------------------
// The pass's events run in one persistent workgroup (k_toi_loop) instead of one host round trip
// per event: per event it takes the smallest cached time of impact over the candidate list (the
// contacts whose cached time is below 1; the list only grows, stale entries are skipped), lists
// the event bodies' contacts from the per-body contact lists, runs the event (toi_event), clears
// the TOI flags of the displaced bodies' contacts, re-synchronises their proxies, recomputes the
// times of impact of the contacts that lost theirs (toi_rescan_wg) and adds those below 1 to the
// list. It returns to the host only when no event is left, when an event re-fattened a proxy
// (the host rebuilds the contact set, as find_new_contacts would add pairs), on a capacity error,
// after max_events events, or, with sub-stepping, after the first solved event (before any time
// of impact is recomputed: the next step's pass recomputes them after its collide). Per event, the
// lists keep the work proportional to the event's own contacts instead of scans over every
// contact or proxy.
constexpr int kLoopBlock = 512;  // (512 threads leave the event code 256 registers per thread)
constexpr int kToiQueued = 4;  // TOI flag: listed for recomputation in this event
enum : int { kLoopDone = 0, kLoopRebuild = 1, kLoopCapacity = 2, kLoopMore = 3, kLoopSubStep = 4, kLoopNeedBig = 5 };

struct LoopArgs {
  const int* cadj_off;  // contacts per non-static body of the current set
  const int* cadj;
  const int* bfix_off;  // fixtures per body
  const int* bfix;
  float4* faabb;
  float4* ffat;
  int* cl;         // event candidates
  int* cl_count;
  int* in_list;    // per contact: in cl
  int* state;      // loop exit state
  int* events;     // [0] events run, [1] events solved
  int max_events;
  int sub_stepping;  // stop after the first solved event
  EventBig* big;     // the global island store (null until an island needed it)
  int force_big;     // every event in the global store (a check of that path)
};

// contacts per non-static body (the adjacency the loop reads)
__global__ void __launch_bounds__(kToiBlock) k_cadj_count(int n, const int2* __restrict__ ab,
                                                          const int* __restrict__ fbody, const int* __restrict__ btype,
                                                          int* __restrict__ deg) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int2 f = ab[k];
  const int x = fbody[f.x], y = fbody[f.y];
  if (btype[x] != 0) atomicAdd(&deg[x], 1);
  if (btype[y] != 0) atomicAdd(&deg[y], 1);
}
__global__ void __launch_bounds__(kToiBlock) k_cadj_fill(int n, const int2* __restrict__ ab,
                                                         const int* __restrict__ fbody, const int* __restrict__ btype,
                                                         const int* __restrict__ off, int* __restrict__ cursor,
                                                         int* __restrict__ adj) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int2 f = ab[k];
  const int x = fbody[f.x], y = fbody[f.y];
  if (btype[x] != 0) adj[off[x] + atomicAdd(&cursor[x], 1)] = k;
  if (btype[y] != 0) adj[off[y] + atomicAdd(&cursor[y], 1)] = k;
}
// fixtures per body
__global__ void __launch_bounds__(kToiBlock) k_bfix_count(int np, const int* __restrict__ fbody, int* __restrict__ deg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) atomicAdd(&deg[fbody[i]], 1);
}
__global__ void __launch_bounds__(kToiBlock) k_bfix_fill(int np, const int* __restrict__ fbody,
                                                         const int* __restrict__ off, int* __restrict__ cursor,
                                                         int* __restrict__ fix) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int b = fbody[i];
  fix[off[b] + atomicAdd(&cursor[b], 1)] = i;
}
// the candidate list from scratch: every contact with a cached time of impact below 1
__global__ void __launch_bounds__(kToiBlock) k_toi_list(ToiArgs a, LoopArgs L) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const bool c = a.tflags[k] == kToiValid && a.tcount[k] <= kMaxSubSteps && a.toi[k] < 1.0f;
  L.in_list[k] = c ? 1 : 0;
  if (c) L.cl[atomicAdd(L.cl_count, 1)] = k;
}

__device__ inline unsigned long long toi_key(float alpha, int k) {
  return (static_cast<unsigned long long>(__float_as_uint(alpha)) << 32) | static_cast<unsigned>(k);
}

__device__ inline unsigned long long wave_min_u64(unsigned long long x) {
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned lo = __shfl_xor(static_cast<unsigned>(x), d, 64);
    const unsigned hi = __shfl_xor(static_cast<unsigned>(x >> 32), d, 64);
    const unsigned long long y = (static_cast<unsigned long long>(hi) << 32) | lo;
    x = y < x ? y : x;
  }
  return x;
}

// contacts of non-static body b whose time of impact is to be recomputed: not cached, a TOI
// candidate, not yet listed in this event
__device__ inline void need_of_body(const ToiArgs& a, const LoopArgs& L, int b) {
  const int e0 = L.cadj_off[b], e1 = L.cadj_off[b + 1];
  for (int e = e0 + threadIdx.x; e < e1; e += kLoopBlock) {
    const int k = L.cadj[e];
    if (a.tflags[k] & (kToiValid | kToiQueued)) continue;
    int bA, bB;
    if (!toi_candidate(a, k, &bA, &bB)) continue;
    if (atomicOr(&a.tflags[k], kToiQueued) & kToiQueued) continue;
    a.need[atomicAdd(&a.flags4[1], 1)] = k;
  }
}

__global__ void __launch_bounds__(kLoopBlock) k_toi_loop(ToiArgs a, LoopArgs L) {
  __shared__ EventShared s;
  __shared__ EventCtl e;
  __shared__ unsigned long long wmin[kLoopBlock / 64];
  __shared__ int stop;
  const int t = threadIdx.x;
  for (int iter = 0;; ++iter) {
    // the smallest (time of impact, contact) over the candidates
    unsigned long long best = ~0ull;
    const int nl = *L.cl_count;
    for (int i = t; i < nl; i += kLoopBlock) {
      const int k = L.cl[i];
      if (a.tflags[k] != kToiValid || a.tcount[k] > kMaxSubSteps) continue;
      const float alpha = a.toi[k];
      if (alpha < 1.0f) {
        const unsigned long long key = toi_key(alpha, k);
        best = key < best ? key : best;
      }
    }
    best = wave_min_u64(best);
    if ((t & 63) == 0) wmin[t >> 6] = best;
    __syncthreads();
    if (t == 0) {
      for (int w = 1; w < kLoopBlock / 64; ++w) best = wmin[w] < best ? wmin[w] : best;
      const float alpha = __uint_as_float(static_cast<unsigned>(best >> 32));
      stop = 0;
      if (best == ~0ull || 1.0f - 10.0f * kFltEpsilon < alpha) stop = kLoopDone + 1;
      else if (iter >= L.max_events) stop = kLoopMore + 1;
      if (stop) *L.state = stop - 1;
      *a.key = best;
      a.flags4[0] = 0;
      a.flags4[1] = 0;
      a.flags4[2] = 0;
      a.flags4[3] = 0;
    }
    __syncthreads();
    if (stop) break;
    // the contacts of the event's dynamic bodies (any order; the event sorts them into set order)
    const unsigned long long key = *a.key;
    const int km = static_cast<int>(key & 0xffffffffu);
    const int2 fm = a.ab[km];
    const int ev[2] = {a.fbody[fm.x], a.fbody[fm.y]};
    // the island has at most 1 + nA + nB contacts and 2 + nA + nB bodies (n: the contacts of a
    // dynamic event body, the only ones gathered from): beyond the LDS store it runs in the
    // global one, which the host allocates on first need (nothing has changed yet at this exit)
    int n_ev[2] = {0, 0};
    for (int side = 0; side < 2; ++side)
      if (a.btype[ev[side]] == 2) n_ev[side] = L.cadj_off[ev[side] + 1] - L.cadj_off[ev[side]];
    const bool big = L.force_big || 2 + n_ev[0] + n_ev[1] > kMaxToiBodies || n_ev[0] > kMaxBodyContacts ||
                     n_ev[1] > kMaxBodyContacts;
    if (big && !L.big) {
      if (t == 0) *L.state = kLoopNeedBig;
      break;
    }
    const int cap = big ? kBigBodyContacts : kMaxBodyContacts;
    for (int side = 0; side < 2; ++side) {
      const int b = ev[side];
      if (a.btype[b] != 2) continue;
      const int e0 = L.cadj_off[b], e1 = L.cadj_off[b + 1];
      for (int i = t; i < e1 - e0 && i < cap; i += kLoopBlock) a.cand[side * cap + i] = L.cadj[e0 + i];
      if (t == 0) a.flags4[2 + side] = e1 - e0;
    }
    __syncthreads();
    if (big) toi_event(a, *L.big, e);
    else toi_event(a, s, e);
    const int status = *a.status;
    if (status == kEventCapacity) {
      if (t == 0) *L.state = kLoopCapacity;
      break;
    }
    // the bodies whose contacts may need a new time of impact: the island's non-static bodies
    // after a solved event, the event's two otherwise (the contact update may have woken them)
    const int* ibody = ev;
    int nbody = 2;
    if (status == kEventSolved) {
      ibody = big ? L.big->bodies : s.bodies;
      nbody = big ? L.big->nbody : s.nbody;
    }
    if (t == 0) {
      L.events[0] += 1;
      if (status == kEventSolved) L.events[1] += 1;
    }
    __syncthreads();
    if (status == kEventSolved) {
      // clear the TOI flags of every contact of the island's dynamic bodies (the event's, and with
      // a bullet the dynamic bodies it gathered; src/private/dynamics/b2_world.rs:871-891), then
      // re-synchronise their proxies
      for (int ib = 0; ib < nbody; ++ib) {
        const int b = ibody[ib];
        if (a.btype[b] != 2) continue;
        const int e0 = L.cadj_off[b], e1 = L.cadj_off[b + 1];
        for (int i = e0 + t; i < e1; i += kLoopBlock) a.tflags[L.cadj[i]] &= ~kToiValid;
        const int f0 = L.bfix_off[b], f1 = L.bfix_off[b + 1];
        for (int i = f0 + t; i < f1; i += kLoopBlock) {
          const int f = L.bfix[i];
          const bool refat = sync_proxy(f, b, a.shapes[f], a.pos0, a.bmass, a.xf, L.faabb, L.ffat);
          a.moved[f] = refat ? 1 : 0;
          if (refat) a.flags4[0] = 1;  // every writer stores 1
        }
      }
      __syncthreads();
      if (a.flags4[0]) {
        // a re-fattened proxy can add pairs: the host rebuilds the contact set first
        if (t == 0) *L.state = kLoopRebuild;
        break;
      }
      if (L.sub_stepping) {
        if (t == 0) *L.state = kLoopSubStep;
        break;
      }
    }
    for (int i = 0; i < nbody; ++i)
      if (a.btype[ibody[i]] != 0) need_of_body(a, L, ibody[i]);
    __syncthreads();
    const int m = a.flags4[1];
    toi_rescan_wg(a);
    __syncthreads();
    for (int i = t; i < m; i += kLoopBlock) {
      const int k = a.rank[i];
      if (a.toi[k] < 1.0f && !L.in_list[k]) {
        L.in_list[k] = 1;
        L.cl[atomicAdd(L.cl_count, 1)] = k;
      }
    }
    __syncthreads();
  }
}

}  // namespace

static ToiArgs toi_args(World* w, float dt, int vi) {
  ContactArrays& c = w->c[w->cur];
  ToiState& ts = w->toi_state[w->cur];
  ToiArgs a;
  a.n = w->n_pairs;
  a.nb = w->nb;
  a.ab = c.ab.p;
  a.hdr = c.hdr.p;
  a.geo = c.geo.p;
  a.pt0 = c.pt0.p;
  a.pt1 = c.pt1.p;
  a.ids = c.ids.p;
  a.toi = ts.toi.p;
  a.tcount = ts.count.p;
  a.tflags = ts.flags.p;
  a.fbody = w->f.body.p;
  a.shapes = w->f.shapev.p;
  a.fmat = w->f.mat.p;
  a.fshape = w->f.shape.p;
  a.btype = w->b.type.p;
  a.bullet = w->b.bullet.p;
  a.pos = w->b.pos.p;
  a.pos0 = w->b.pos0.p;
  a.vel = w->b.vel.p;
  a.xf = w->b.xf.p;
  a.bmass = w->b.mass.p;
  a.alpha0 = w->toi_alpha0.p;
  a.awake = w->sleep_possible() ? w->b.awake.p : nullptr;
  a.sleep_time = w->b.sleep_time.p;
  a.key = w->toi_ctl.p;
  a.status = reinterpret_cast<int*>(w->toi_ctl.p + 1);
  a.event_bodies = reinterpret_cast<int2*>(w->toi_ctl.p + 2);
  a.flags4 = reinterpret_cast<int*>(w->toi_ctl.p + 3);
  a.cand = w->toi_cand.p;
  a.need = w->toi_need.p;
  a.rank = w->toi_rank.p;
  a.snap = w->toi_snap.p;
  a.moved = w->f.moved.p;
  a.np = w->np;
  a.block_solve = w->block_solve ? 1 : 0;
  a.velocity_iterations = vi;
  a.dt = dt;
  return a;
}

static int toi_reserve(World* w) {
  int st;
  const size_t n = pair_capacity(w, w->n_pairs);
  ToiState& ts = w->toi_state[w->cur];
  if ((st = ts.toi.reserve(n)) || (st = ts.count.reserve(n)) || (st = ts.flags.reserve(n))) return st;
  if ((st = w->toi_need.reserve(n)) || (st = w->toi_rank.reserve(n)) || (st = w->toi_snap.reserve(19 * n)) ||
      (st = w->toi_cand.reserve(2 * kBigBodyContacts)))
    return st;
  if ((st = w->toi_alpha0.reserve(std::max(w->nb, 1))) || (st = w->toi_ctl.reserve(7)) ||
      (st = w->toi_cl.reserve(n)) || (st = w->toi_in_list.reserve(n)))
    return st;
  return B2G_OK;
}

// ctl: [0] key, [1] status and candidate count, [2] event bodies, [3] / [4] flags4, [5] loop
// state and events run, [6] events solved
struct ToiCtl {
  unsigned long long key;
  int status, cl_count;
  int bodies[2];
  int flags4[4];
  int loop_state, events, solved, pad;
};
static_assert(sizeof(ToiCtl) == 7 * sizeof(unsigned long long), "TOI control block");

static int toi_read(World* w, ToiCtl* out) {
  B2G_HIP(hipMemcpyAsync(out, w->toi_ctl.p, sizeof(ToiCtl), hipMemcpyDeviceToHost, w->stream));
  B2G_HIP(hipStreamSynchronize(w->stream));
  return B2G_OK;
}

int carry_manifolds_for_toi(World* w);

// find_new_contacts: a new contact set from the fat AABBs, its manifolds and TOI state carried
__global__ void __launch_bounds__(kToiBlock) k_clear(int n, int* __restrict__ a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = 0;
}

// Touching contacts of the predecessor set with no successor in the current set: the next collide
// destroys them and wakes their bodies (src/private/dynamics/b2_contact.rs:33-45), and
// b2g_world_get_destroyed_contacts reports them. A TOI rebuild makes the current set the
// predecessor and overwrites the old predecessor, so these entries are gathered first and appended
// to the new predecessor after its own pairs (where nothing takes them as a predecessor).
__global__ void __launch_bounds__(kToiBlock) k_mark_successors(int n, const int* __restrict__ old_index, int n_old,
                                                               int* __restrict__ succ) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int o = old_index[k];
  if (o >= 0 && o < n_old) succ[o] = 1;
}

__global__ void __launch_bounds__(kToiBlock) k_dropped_flags(int n_old, const int* __restrict__ hdr,
                                                             const int* __restrict__ succ, int* __restrict__ flag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_old) return;
  flag[k] = (hdr[k] >> 8) > 0 && !succ[k] ? 1 : 0;
}

struct ContactView {
  int2* pair;
  int2* ab;
  int* hdr;
  float4 *geo, *pt0, *pt1;
  uint2* ids;
  int* colour;
};

__global__ void __launch_bounds__(kToiBlock) k_move_dropped(int n, const int* __restrict__ flag,
                                                            const int* __restrict__ pos, ContactView src,
                                                            ContactView dst, int dst_off) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || (flag && !flag[k])) return;
  const int d = dst_off + (pos ? pos[k] : k);
  dst.pair[d] = src.pair[k];
  dst.ab[d] = src.ab[k];
  dst.hdr[d] = src.hdr[k];
  dst.geo[d] = src.geo[k];
  dst.pt0[d] = src.pt0[k];
  dst.pt1[d] = src.pt1[k];
  dst.ids[d] = src.ids[k];
  if (dst.colour) dst.colour[d] = -1;
}

static ContactView view_of(ContactArrays& c) {
  return ContactView{c.pair.p, c.ab.p, c.hdr.p, c.geo.p, c.pt0.p, c.pt1.p, c.ids.p, c.colour.p};
}

// the flagged entries of src (flag and pos its exclusive scan; null: all n) copied densely to dst
// from dst_off
static int move_dropped(World* w, int n, const int* flag, const int* pos, ContactArrays& src, ContactArrays& dst,
                        int dst_off) {
  hipLaunchKernelGGL(k_move_dropped, dim3(blocks_for(n, kToiBlock)), dim3(kToiBlock), 0, w->stream, n, flag, pos,
                     view_of(src), view_of(dst), dst_off);
  B2G_HIP(hipGetLastError());
  return B2G_OK;
}

static int reserve_contacts(ContactArrays& c, int n, int keep) {
  const size_t m = std::max(n, 1), k = std::max(keep, 0);
  int st;
  if ((st = c.pair.reserve(m, true, k)) || (st = c.ab.reserve(m, true, k)) || (st = c.hdr.reserve(m, true, k)) ||
      (st = c.geo.reserve(m, true, k)) || (st = c.pt0.reserve(m, true, k)) || (st = c.pt1.reserve(m, true, k)) ||
      (st = c.ids.reserve(m, true, k)) || (st = c.colour.reserve(m, true, k)))
    return st;
  return B2G_OK;
}

static int toi_rebuild(World* w) {
  int st;
  // the predecessor's dropped touching contacts (its extra entries included) before it is
  // overwritten; when the current set has been through a collide already (a sub-stepped pass
  // resumed after the step's collide), that collide destroyed them
  const int n_old = w->n_pairs_old, n_cur = w->n_pairs;
  int dropped = 0;
  if (w->collide_pending && n_old > 0 && w->manifolds_valid[1 - w->cur]) {
    ContactArrays& pred = w->c[1 - w->cur];
    if ((st = w->toi_succ.reserve(n_old)) || (st = w->toi_drop_flag.reserve(n_old)) ||
        (st = w->toi_drop_pos.reserve(n_old)) || (st = w->toi_drop_total.reserve(1)))
      return st;
    hipLaunchKernelGGL(k_clear, dim3(blocks_for(n_old, kToiBlock)), dim3(kToiBlock), 0, w->stream, n_old,
                       w->toi_succ.p);
    if (n_cur > 0)
      hipLaunchKernelGGL(k_mark_successors, dim3(blocks_for(n_cur, kToiBlock)), dim3(kToiBlock), 0, w->stream, n_cur,
                         w->c[w->cur].old_index.p, n_old, w->toi_succ.p);
    hipLaunchKernelGGL(k_dropped_flags, dim3(blocks_for(n_old, kToiBlock)), dim3(kToiBlock), 0, w->stream, n_old,
                       pred.hdr.p, w->toi_succ.p, w->toi_drop_flag.p);
    B2G_HIP(hipGetLastError());
    if ((st = exclusive_scan(w, w->toi_drop_flag.p, w->toi_drop_pos.p, n_old, w->toi_drop_total.p))) return st;
    B2G_HIP(hipMemcpyAsync(&dropped, w->toi_drop_total.p, sizeof(int), hipMemcpyDeviceToHost, w->stream));
    B2G_HIP(hipStreamSynchronize(w->stream));
    if (dropped > 0) {
      if ((st = reserve_contacts(w->toi_dropped, dropped, 0))) return st;
      if ((st = move_dropped(w, n_old, w->toi_drop_flag.p, w->toi_drop_pos.p, pred, w->toi_dropped, 0))) return st;
    }
  }
  if ((st = phase_broadphase(w))) return st;
  if (dropped > 0) {
    // appended to the new predecessor (the set the TOI pass started from) after its own pairs
    ContactArrays& pred = w->c[1 - w->cur];
    const int base = w->n_pairs_old;
    if ((st = reserve_contacts(pred, base + dropped, base))) return st;
    if ((st = move_dropped(w, dropped, nullptr, nullptr, w->toi_dropped, pred, base))) return st;
    w->n_pairs_old = base + dropped;
  }
  w->manifolds_valid[w->cur] = false;
  w->collide_pending = true;
  if ((st = carry_manifolds_for_toi(w))) return st;
  return toi_carry_state(w);
}

// fixtures per body (rebuilt when the fixture count changed)
static int toi_body_fixtures(World* w) {
  if (w->toi_bfix_np == w->np) return B2G_OK;
  int st;
  const int nb = w->nb, np = w->np;
  if ((st = w->toi_bfix_off.reserve(nb + 1)) || (st = w->toi_bfix_cur.reserve(nb + 1)) ||
      (st = w->toi_bfix.reserve(std::max(np, 1))))
    return st;
  hipLaunchKernelGGL(k_clear, dim3(blocks_for(nb + 1, kToiBlock)), dim3(kToiBlock), 0, w->stream, nb + 1,
                     w->toi_bfix_cur.p);
  if (np > 0)
    hipLaunchKernelGGL(k_bfix_count, dim3(blocks_for(np, kToiBlock)), dim3(kToiBlock), 0, w->stream, np, w->f.body.p,
                       w->toi_bfix_cur.p);
  if ((st = exclusive_scan(w, w->toi_bfix_cur.p, w->toi_bfix_off.p, nb + 1, nullptr))) return st;
  hipLaunchKernelGGL(k_clear, dim3(blocks_for(nb + 1, kToiBlock)), dim3(kToiBlock), 0, w->stream, nb + 1,
                     w->toi_bfix_cur.p);
  if (np > 0)
    hipLaunchKernelGGL(k_bfix_fill, dim3(blocks_for(np, kToiBlock)), dim3(kToiBlock), 0, w->stream, np, w->f.body.p,
                       w->toi_bfix_off.p, w->toi_bfix_cur.p, w->toi_bfix.p);
  B2G_HIP(hipGetLastError());
  w->toi_bfix_np = np;
  return B2G_OK;
}

static LoopArgs loop_args(World* w) {
  LoopArgs L;
  L.cadj_off = w->toi_cadj_off.p;
  L.cadj = w->toi_cadj.p;
  L.bfix_off = w->toi_bfix_off.p;
  L.bfix = w->toi_bfix.p;
  L.faabb = w->f.aabb.p;
  L.ffat = w->f.fat.p;
  L.cl = w->toi_cl.p;
  int* ctl = reinterpret_cast<int*>(w->toi_ctl.p);
  L.cl_count = ctl + 3;
  L.in_list = w->toi_in_list.p;
  L.state = ctl + 10;
  L.events = ctl + 11;
  L.max_events = 1 << 16;
  L.sub_stepping = w->sub_stepping ? 1 : 0;
  L.big = w->toi_big.p ? reinterpret_cast<EventBig*>(w->toi_big.p) : nullptr;
  L.force_big = w->toi_force_big ? 1 : 0;
  return L;
}

// the loop's per-set structures for the current contact set: contacts per non-static body and
// the candidate list
static int toi_loop_setup(World* w, float dt, int vi) {
  int st;
  const int nb = w->nb, n = w->n_pairs;
  if ((st = w->toi_cadj_off.reserve(nb + 1)) || (st = w->toi_cadj_cur.reserve(nb + 1)) ||
      (st = w->toi_cadj.reserve(std::max(2 * n, 1))) || (st = w->toi_cl.reserve(std::max(n, 1))) ||
      (st = w->toi_in_list.reserve(std::max(n, 1))))
    return st;
  hipLaunchKernelGGL(k_clear, dim3(blocks_for(nb + 1, kToiBlock)), dim3(kToiBlock), 0, w->stream, nb + 1,
                     w->toi_cadj_cur.p);
  const ToiArgs a = toi_args(w, dt, vi);
  if (n > 0)
    hipLaunchKernelGGL(k_cadj_count, dim3(blocks_for(n, kToiBlock)), dim3(kToiBlock), 0, w->stream, n, a.ab, a.fbody,
                       a.btype, w->toi_cadj_cur.p);
  if ((st = exclusive_scan(w, w->toi_cadj_cur.p, w->toi_cadj_off.p, nb + 1, nullptr))) return st;
  hipLaunchKernelGGL(k_clear, dim3(blocks_for(nb + 1, kToiBlock)), dim3(kToiBlock), 0, w->stream, nb + 1,
                     w->toi_cadj_cur.p);
  const LoopArgs L = loop_args(w);
  hipLaunchKernelGGL(k_clear, dim3(1), dim3(kToiBlock), 0, w->stream, 1, L.cl_count);
  if (n > 0) {
    hipLaunchKernelGGL(k_cadj_fill, dim3(blocks_for(n, kToiBlock)), dim3(kToiBlock), 0, w->stream, n, a.ab, a.fbody,
                       a.btype, w->toi_cadj_off.p, w->toi_cadj_cur.p, w->toi_cadj.p);
    hipLaunchKernelGGL(k_toi_list, dim3(blocks_for(n, kToiBlock)), dim3(kToiBlock), 0, w->stream, a, L);
  }
  B2G_HIP(hipGetLastError());
  return B2G_OK;
}

// The TOI state of a rebuilt contact set: each pair's cached time of impact, TOI count and flags
// from its predecessor in the previous set (a new pair has none: its time is computed when the
// pass next scans it)
int toi_carry_state(World* w) {
  int st;
  if ((st = toi_reserve(w))) return st;
  if (w->n_pairs > 0) {
    ToiState& o = w->toi_state[1 - w->cur];
    ToiState& c = w->toi_state[w->cur];
    hipLaunchKernelGGL(k_toi_carry, dim3(blocks_for(w->n_pairs, kToiBlock)), dim3(kToiBlock), 0, w->stream,
                       w->n_pairs, w->c[w->cur].old_index.p, o.toi.p, o.count.p, o.flags.p, c.toi.p, c.count.p,
                       c.flags.p);
    B2G_HIP(hipGetLastError());
  }
  return B2G_OK;
}

// The TOI pass after the solve (src/private/dynamics/b2_world.rs:534-901). The current contact
// set is the one find_new_contacts left. The events run on the device (k_toi_loop); the host
// rebuilds the contact set when an event re-fattened a proxy. A pass starts from alpha0 = 0 and no
// cached times when the last pass was complete; with sub-stepping a pass ends after its first
// solved event and the next step's pass resumes it (step_complete).
int phase_toi(World* w, float dt, int vi) {
  int st;
  w->toi_events = 0;
  const bool resume = !w->step_complete;
  if ((st = carry_manifolds_for_toi(w))) return st;
  if ((st = toi_reserve(w))) return st;
  if ((st = toi_body_fixtures(w))) return st;
  static const ToiCtl kReset = {~0ull, 0, 0, {0, 0}, {0, 0, 0, 0}, 0, 0, 0, 0};
  B2G_HIP(hipMemcpyAsync(w->toi_ctl.p, &kReset, sizeof(kReset), hipMemcpyHostToDevice, w->stream));
  if (w->n_pairs > 0) {
    const ToiArgs b = toi_args(w, dt, vi);
    if (!resume) {
      hipLaunchKernelGGL(k_toi_first, dim3(blocks_for(w->n_pairs, kToiBlock)), dim3(kToiBlock), 0, w->stream, b);
    } else {
      // the times the last sub-step's event invalidated, and those of pairs found since, in set
      // order with the sweep alignment (the reference's scan recomputes them in list order)
      hipLaunchKernelGGL(k_toi_resume, dim3(blocks_for(w->n_pairs, kToiBlock)), dim3(kToiBlock), 0, w->stream,
                         w->n_pairs, b.tflags);
      hipLaunchKernelGGL(k_toi_need, dim3(blocks_for(w->n_pairs, kToiBlock)), dim3(kToiBlock), 0, w->stream, b, 0);
      hipLaunchKernelGGL(k_toi_rescan, dim3(1), dim3(kRescanBlock), 0, w->stream, b);
    }
    B2G_HIP(hipGetLastError());
  }
  if ((st = toi_loop_setup(w, dt, vi))) return st;
  bool dirty = false;
  long long total = 0;
  // every event either disables its contact or raises its TOI count, so the pass ends; the bound
  // is a guard against a logic error, not a limit of the reference
  const long long kMaxEvents = 4ll * (kMaxSubSteps + 2) * (static_cast<long long>(w->n_pairs) + w->nb + 16);
  for (;;) {
    if (w->n_pairs == 0) {
      w->step_complete = true;
      return dirty ? toi_rebuild(w) : B2G_OK;
    }
    hipLaunchKernelGGL(k_toi_loop, dim3(1), dim3(kLoopBlock), 0, w->stream, toi_args(w, dt, vi), loop_args(w));
    B2G_HIP(hipGetLastError());
    ToiCtl ctl;
    if ((st = toi_read(w, &ctl))) return st;
    w->toi_events += ctl.solved;
    total += ctl.events;
    if (ctl.events > 0) dirty = true;
    if (ctl.loop_state == kLoopCapacity) {
      set_error("continuous collision: a TOI island has more than 4096 bodies, 4096 contacts or 16384 contacts of "
                "one body");
      return B2G_ERR_CAPACITY;
    }
    if (total > kMaxEvents) {
      set_error("continuous collision: the TOI pass did not end");
      return B2G_ERR_DEVICE;
    }
    // the next launch counts its own events
    static const int kZero[3] = {kLoopDone, 0, 0};
    B2G_HIP(hipMemcpyAsync(reinterpret_cast<int*>(w->toi_ctl.p) + 10, kZero, sizeof(kZero), hipMemcpyHostToDevice,
                           w->stream));
    if (ctl.loop_state == kLoopNeedBig) {
      // the global island store, once per world; the loop runs the same event again with it
      if (w->toi_big.p) {
        set_error("continuous collision: the island store was not taken");
        return B2G_ERR_DEVICE;
      }
      if ((st = w->toi_big.reserve(sizeof(EventBig)))) return st;
      continue;
    }
    if (ctl.loop_state == kLoopDone || ctl.loop_state == kLoopSubStep) {
      // a pass without another event is complete (src/private/dynamics/b2_world.rs:676-679); a
      // sub-step ends after its solved event (:896-899)
      w->step_complete = ctl.loop_state == kLoopDone;
      // manifolds that an event changed after the last rebuild become the next collide's
      // predecessors through one more rebuild
      return dirty ? toi_rebuild(w) : B2G_OK;
    }
    if (ctl.loop_state == kLoopRebuild) {
      // find_new_contacts: only re-fattened proxies can add pairs (the reference queries the
      // moved proxies only), so the set is rebuilt only when a proxy was re-fattened
      if ((st = toi_rebuild(w))) return st;
      dirty = false;
      if (w->sub_stepping) {
        // the rebuild followed a solved event: the sub-step ends here
        w->step_complete = false;
        return B2G_OK;
      }
      if (w->n_pairs > 0) {
        const ToiArgs b = toi_args(w, dt, vi);
        hipLaunchKernelGGL(k_clear, dim3(1), dim3(kToiBlock), 0, w->stream, 1, b.flags4 + 1);
        hipLaunchKernelGGL(k_toi_need, dim3(blocks_for(w->n_pairs, kToiBlock)), dim3(kToiBlock), 0, w->stream, b, 0);
        hipLaunchKernelGGL(k_toi_rescan, dim3(1), dim3(kRescanBlock), 0, w->stream, b);
        B2G_HIP(hipGetLastError());
        if ((st = toi_loop_setup(w, dt, vi))) return st;
      }
    }
  }
}

}  // namespace b2g
