// b2g narrowphase: B2contactManager::collide + B2contact::update for every contact of the
// current set, one thread per contact (src/private/dynamics/b2_contact_manager.rs:83-171,
// src/private/dynamics/b2_contact.rs:103-217).
//
// The contact set is exactly the set of fat-AABB-overlapping pairs (see b2g_broadphase.hip),
// so the reference's "destroy if the fat AABBs no longer overlap" branch never fires here.
// Each thread evaluates the manifold with the bodies' current transforms, then copies the
// normal/tangent impulses of the previous manifold of the same pair by contact-feature id.
// Manifolds are stored SoA (header, geometry, two points, ids) so the writes of a wave are
// coalesced.
#include <mutex>

#include "b2g_internal.h"
#include "b2g_distance.h"

namespace b2g {

static constexpr int kBlock = 128;

__device__ inline void store_manifold(int k, const Manifold& m, const float imp[4], int* hdr, float4* geo,
                                      float4* pt0, float4* pt1, uint2* ids) {
  hdr[k] = m.type | (m.count << 8);
  geo[k] = make_float4(m.local_normal.x, m.local_normal.y, m.local_point.x, m.local_point.y);
  pt0[k] = make_float4(m.lp[0].x, m.lp[0].y, imp[0], imp[1]);
  pt1[k] = make_float4(m.lp[1].x, m.lp[1].y, imp[2], imp[3]);
  ids[k] = make_uint2(m.id[0], m.id[1]);
}

// Sleep (when the world can have sleeping bodies: awake non-null). A contact none of whose bodies
// is active (awake and not static) is not updated: it keeps its manifold and impulses
// (src/private/dynamics/b2_contact_manager.rs:104-114). A contact whose touching state changes
// asks to wake both bodies (src/private/dynamics/b2_contact.rs:203-206); the requests are applied
// after the pass, so whether a contact is updated depends on the awake states at the start of
// the pass (the reference's sequential pass lets a body woken by an earlier contact of its list
// update later ones).
struct SleepView {
  const int* awake;     // null: every non-static body is awake
  const int* btype;
  int* wake;
  const float4* ogeo;   // previous manifold geometry (copied for contacts that are not updated)
};

// The shape-pair kinds of a step's pairs are interleaved (the pair list is in tree order), and a
// wave holding several kinds runs each kind's code in turn. So a workgroup first sorts its pairs
// by kind in LDS (a counting sort over the 16 shape-type pairs) and each thread then evaluates
// the pair at its sorted position: most waves hold one kind. A thread's results go to the pair's
// slot in an LDS copy of the workgroup's output range, which is then written out in pair order
// (coalesced rows), so the thread that evaluates a pair does not change any result.
static constexpr int kCollideBlock = 256;
static constexpr int kKinds = 17;  // 4 x 4 shape-type pairs (lower type first), and past-the-end

struct CollideStage {
  float4 geo[kCollideBlock], pt0[kCollideBlock], pt1[kCollideBlock];
  int2 ab[kCollideBlock];
  uint2 ids[kCollideBlock];
  int hdr[kCollideBlock], touching[kCollideBlock];
};

__global__ void __launch_bounds__(kCollideBlock) k_collide(int n, const int2* __restrict__ pairs,
                                                    const int* __restrict__ old_index, const int* __restrict__ fbody,
                                                    const int4* __restrict__ fshape, const ShapeV* __restrict__ shapes,
                                                    const float4* __restrict__ bxf,
                                                    // previous manifolds
                                                    const int* __restrict__ ohdr, const float4* __restrict__ opt0,
                                                    const float4* __restrict__ opt1, const uint2* __restrict__ oids,
                                                    SleepView sv,
                                                    // outputs
                                                    int2* __restrict__ ab, int* __restrict__ hdr,
                                                    float4* __restrict__ geo, float4* __restrict__ pt0,
                                                    float4* __restrict__ pt1, uint2* __restrict__ ids,
                                                    int* __restrict__ touching) {
  __shared__ int kind_count[kKinds];
  __shared__ int sorted[kCollideBlock];
  __shared__ CollideStage st;
  const int t = threadIdx.x;
  const int base = blockIdx.x * kCollideBlock;
  if (t < kKinds) kind_count[t] = 0;
  __syncthreads();
  int kind = kKinds - 1;
  if (base + t < n) {
    const int2 p = pairs[base + t];
    const int ta = fshape[p.x].x & 3, tb = fshape[p.y].x & 3;
    kind = ta < tb ? ta * 4 + tb : tb * 4 + ta;
  }
  const int r = atomicAdd(&kind_count[kind], 1);
  __syncthreads();
  int pos = r;
  for (int j = 0; j < kind; ++j) pos += kind_count[j];
  sorted[pos] = t;
  __syncthreads();
  const int s = sorted[t];  // the pair this thread evaluates: base + s
  const int k = base + s;
  if (k < n) {
    const int2 pr = pairs[k];
    // the reference pair is (lower proxy id, higher proxy id); b2_contact_create then lets the
    // registry decide which fixture is A
    const int lo = pr.x < pr.y ? pr.x : pr.y, hi = pr.x < pr.y ? pr.y : pr.x;
    const int4 hq = fshape[lo], hp = fshape[hi];
    const bool swap = registry_swap(hq.x & 0xff, hp.x & 0xff) == 1;
    const int fa = swap ? hi : lo, fb = swap ? lo : hi;
    const int bA = fbody[fa], bB = fbody[fb];
    const int oi = old_index[k];
    st.ab[s] = make_int2(fa, fb);
    bool frozen = false;
    if (sv.awake) {
      const bool active_a = sv.btype[bA] != 0 && sv.awake[bA], active_b = sv.btype[bB] != 0 && sv.awake[bB];
      frozen = !active_a && !active_b;
    }
    if (frozen) {
      if (oi >= 0) {
        const int oh = ohdr[oi];
        st.hdr[s] = oh;
        st.geo[s] = sv.ogeo[oi];
        st.pt0[s] = opt0[oi];
        st.pt1[s] = opt1[oi];
        st.ids[s] = oids[oi];
        st.touching[s] = (oh >> 8) > 0 ? 1 : 0;
      } else {
        const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        st.hdr[s] = 0;
        st.geo[s] = z;
        st.pt0[s] = z;
        st.pt1[s] = z;
        st.ids[s] = make_uint2(0u, 0u);
        st.touching[s] = 0;
      }
    } else if (hq.w | hp.w) {
      // a sensor: no manifold, no wake (src/private/dynamics/b2_contact.rs:148-162); its overlap
      // comes from k_sensor_overlap (the GJK distance would give this kernel a scratch frame)
      const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      st.hdr[s] = kSensorPending;
      st.geo[s] = z;
      st.pt0[s] = z;
      st.pt1[s] = z;
      st.ids[s] = make_uint2(0u, 0u);
      st.touching[s] = 0;
    } else {
      // shapes are read in place from their records (no private copy of the vertex arrays)
      const ShapeV& sa = shapes[fa];
      const ShapeV& sb = shapes[fb];
      Xf xa = xf_of(bxf[bA]);
      Xf xb = xf_of(bxf[bB]);
      Manifold m;
      evaluate(&m, sa, xa, sb, xb);
      float imp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      const int ocount = oi >= 0 ? ohdr[oi] >> 8 : 0;
      if (m.count > 0 && ocount > 0) {
        const uint2 oid = oids[oi];
        const float4 o0 = opt0[oi], o1 = opt1[oi];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          if (i >= m.count) break;
          if (m.id[i] == oid.x) { imp[2 * i] = o0.z; imp[2 * i + 1] = o0.w; }
          else if (ocount > 1 && m.id[i] == oid.y) { imp[2 * i] = o1.z; imp[2 * i + 1] = o1.w; }
        }
      }
      store_manifold(s, m, imp, st.hdr, st.geo, st.pt0, st.pt1, st.ids);
      st.touching[s] = m.count > 0 ? 1 : 0;
      if (sv.awake && (m.count > 0) != (ocount > 0)) {
        if (sv.btype[bA] != 0) sv.wake[bA] = 1;
        if (sv.btype[bB] != 0) sv.wake[bB] = 1;
      }
    }
  }
  __syncthreads();
  const int ko = base + t;
  if (ko < n) {
    ab[ko] = st.ab[t];
    hdr[ko] = st.hdr[t];
    geo[ko] = st.geo[t];
    pt0[ko] = st.pt0[t];
    pt1[ko] = st.pt1[t];
    ids[ko] = st.ids[t];
    touching[ko] = st.touching[t];
  }
}

// b2_contact_destroy wakes both bodies of a touching contact (src/private/dynamics/b2_contact.rs:
// 33-45); the collide pass destroys a contact with an active body whose fat AABBs stopped
// overlapping (src/private/dynamics/b2_contact_manager.rs:117-127). Those are the previous set's
// pairs that are not in the current set.
__global__ void __launch_bounds__(kBlock) k_wake_destroyed(int n_old, const int2* __restrict__ old_pairs,
                                                           const int* __restrict__ ohdr, const int* __restrict__ fbody,
                                                           const float4* __restrict__ fat, SleepView sv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_old || (ohdr[k] >> 8) == 0) return;
  const int2 pr = old_pairs[k];
  const int bA = fbody[pr.x], bB = fbody[pr.y];
  const bool active_a = sv.btype[bA] != 0 && sv.awake[bA], active_b = sv.btype[bB] != 0 && sv.awake[bB];
  if (!active_a && !active_b) return;
  const float4 fa = fat[pr.x], fb = fat[pr.y];
  Aabb a, b;
  a.lo = v2(fa.x, fa.y);
  a.hi = v2(fa.z, fa.w);
  b.lo = v2(fb.x, fb.y);
  b.hi = v2(fb.z, fb.w);
  if (overlap(a, b)) return;
  if (sv.btype[bA] != 0) sv.wake[bA] = 1;
  if (sv.btype[bB] != 0) sv.wake[bB] = 1;
}

// The sensor contacts the collide updated: b2_test_overlap of their shapes
// (src/private/collision/b2_collision.rs:216-245) as the touching flag
__global__ void __launch_bounds__(kBlock) k_sensor_overlap(int n, const int2* __restrict__ ab,
                                                           const int* __restrict__ fbody,
                                                           const ShapeV* __restrict__ shapes,
                                                           const float4* __restrict__ bxf, int* __restrict__ hdr) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || hdr[k] != kSensorPending) return;
  const int2 f = ab[k];
  const bool overlap = test_overlap(shapes[f.x], xf_of(bxf[fbody[f.x]]), shapes[f.y], xf_of(bxf[fbody[f.y]]));
  hdr[k] = overlap ? kSensorTouching : 0;
}

// B2body::set_awake(true) for the bodies asked to wake (src/private/dynamics/b2_body.rs:783-801)
__global__ void __launch_bounds__(kBlock) k_apply_wakes(int nb, int* __restrict__ wake, int* __restrict__ awake,
                                                        float* __restrict__ sleep_time) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb || !wake[b]) return;
  wake[b] = 0;
  awake[b] = 1;
  sleep_time[b] = 0.0f;
}

int phase_collide(World* w) {
  int n = w->n_pairs;
  int st;
  if ((st = w->s.touch_flag.reserve(pair_capacity(w, n)))) return st;
  ContactArrays& c = w->c[w->cur];
  ContactArrays& o = w->c[1 - w->cur];
  const bool sleepy = w->sleep_possible();
  SleepView sv{sleepy ? w->b.awake.p : nullptr, w->b.type.p, w->b.wake.p, o.geo.p};
  if (sleepy && w->n_pairs_old > 0)
    hipLaunchKernelGGL(k_wake_destroyed, dim3(blocks_for(w->n_pairs_old, kBlock)), dim3(kBlock), 0, w->stream,
                       w->n_pairs_old, o.pair.p, o.hdr.p, w->f.body.p, w->f.fat.p, sv);
  if (n > 0)
    launch_timed(w, B2G_KERNEL_COLLIDE, n, k_collide, dim3(blocks_for(n, kCollideBlock)), dim3(kCollideBlock), n, c.pair.p,
                 c.old_index.p, w->f.body.p, w->f.shape.p, w->f.shapev.p, w->b.xf.p, o.hdr.p, o.pt0.p, o.pt1.p,
                 o.ids.p, sv, c.ab.p, c.hdr.p, c.geo.p, c.pt0.p, c.pt1.p, c.ids.p, w->s.touch_flag.p);
  if (w->n_sensors > 0 && n > 0)
    hipLaunchKernelGGL(k_sensor_overlap, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, w->stream, n, c.ab.p,
                       w->f.body.p, w->f.shapev.p, w->b.xf.p, c.hdr.p);
  if (sleepy && w->nb > 0)
    hipLaunchKernelGGL(k_apply_wakes, dim3(blocks_for(w->nb, kBlock)), dim3(kBlock), 0, w->stream, w->nb,
                       w->b.wake.p, w->b.awake.p, w->b.sleep_time.p);
  B2G_HIP(hipGetLastError());
  return B2G_OK;
}

// ---- stateless entry point ----------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_collide_given(int n, const int2* __restrict__ pairs,
                                                          const float4* __restrict__ xa4, const float4* __restrict__ xb4,
                                                          const ShapeV* __restrict__ shapes, int* __restrict__ hdr,
                                                          float4* __restrict__ geo, float4* __restrict__ pt0,
                                                          float4* __restrict__ pt1, uint2* __restrict__ ids) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int2 pr = pairs[k];
  const ShapeV& sa = shapes[pr.x];
  const ShapeV& sb = shapes[pr.y];
  Manifold m;
  evaluate(&m, sa, xf_of(xa4[k]), sb, xf_of(xb4[k]));
  float imp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  store_manifold(k, m, imp, hdr, geo, pt0, pt1, ids);
}

// the fixture indices are in range and every pair is in registry order
static int check_given_pairs(const World* w, int n, const int32_t* pairs, const char* who) {
  for (int i = 0; i < 2 * n; ++i)
    if (pairs[i] < 0 || pairs[i] >= w->np) {
      set_error(std::string(who) + ": fixture index out of range");
      return B2G_ERR_INVALID_ARGUMENT;
    }
  for (int i = 0; i < n; ++i) {
    int ta = w->h_fshape[pairs[2 * i]].x & 0xff, tb = w->h_fshape[pairs[2 * i + 1]].x & 0xff;
    if (registry_swap(ta, tb) != 0) {
      set_error(std::string(who) + ": pair is not in registry order (A must be the primary shape type)");
      return B2G_ERR_INVALID_ARGUMENT;
    }
  }
  return B2G_OK;
}

int collide_fixture_pairs(World* w, int n, const int32_t* pairs, const float* xfa, const float* xfb,
                          b2g_manifold* out) {
  if (n <= 0) return B2G_OK;
  int st0 = check_given_pairs(w, n, pairs, "b2g_collide_pairs");
  if (st0) return st0;
  DevArray<int2> dp;
  DevArray<float4> da, db, geo, p0, p1;
  DevArray<int> hdr;
  DevArray<uint2> ids;
  int st;
  if ((st = dp.reserve(n)) || (st = da.reserve(n)) || (st = db.reserve(n)) || (st = geo.reserve(n)) ||
      (st = p0.reserve(n)) || (st = p1.reserve(n)) || (st = hdr.reserve(n)) || (st = ids.reserve(n)))
    return st;
  std::vector<int> hh(n);
  std::vector<float4> hg(n), h0(n), h1(n);
  std::vector<uint2> hi(n);
  bool ok = hipMemcpyAsync(dp.p, pairs, sizeof(int2) * n, hipMemcpyHostToDevice, w->stream) == hipSuccess &&
            hipMemcpyAsync(da.p, xfa, sizeof(float4) * n, hipMemcpyHostToDevice, w->stream) == hipSuccess &&
            hipMemcpyAsync(db.p, xfb, sizeof(float4) * n, hipMemcpyHostToDevice, w->stream) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(k_collide_given, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, w->stream, n, dp.p, da.p, db.p,
                       w->f.shapev.p, hdr.p, geo.p, p0.p, p1.p, ids.p);
    ok = hipGetLastError() == hipSuccess &&
         hipMemcpyAsync(hh.data(), hdr.p, sizeof(int) * n, hipMemcpyDeviceToHost, w->stream) == hipSuccess &&
         hipMemcpyAsync(hg.data(), geo.p, sizeof(float4) * n, hipMemcpyDeviceToHost, w->stream) == hipSuccess &&
         hipMemcpyAsync(h0.data(), p0.p, sizeof(float4) * n, hipMemcpyDeviceToHost, w->stream) == hipSuccess &&
         hipMemcpyAsync(h1.data(), p1.p, sizeof(float4) * n, hipMemcpyDeviceToHost, w->stream) == hipSuccess &&
         hipMemcpyAsync(hi.data(), ids.p, sizeof(uint2) * n, hipMemcpyDeviceToHost, w->stream) == hipSuccess &&
         hipStreamSynchronize(w->stream) == hipSuccess;
  }
  dp.release(); da.release(); db.release(); geo.release(); p0.release(); p1.release(); hdr.release(); ids.release();
  if (!ok) {
    set_error("b2g_collide_pairs: device error");
    return B2G_ERR_DEVICE;
  }
  for (int i = 0; i < n; ++i) {
    b2g_manifold& m = out[i];
    m.type = hh[i] & 0xff;
    m.point_count = hh[i] >> 8;
    m.local_normal[0] = hg[i].x; m.local_normal[1] = hg[i].y;
    m.local_point[0] = hg[i].z; m.local_point[1] = hg[i].w;
    m.points[0].local_point[0] = h0[i].x; m.points[0].local_point[1] = h0[i].y;
    m.points[0].normal_impulse = h0[i].z; m.points[0].tangent_impulse = h0[i].w;
    m.points[0].id = hi[i].x; m.points[0].pad = 0;
    m.points[1].local_point[0] = h1[i].x; m.points[1].local_point[1] = h1[i].y;
    m.points[1].normal_impulse = h1[i].z; m.points[1].tangent_impulse = h1[i].w;
    m.points[1].id = hi[i].y; m.points[1].pad = 0;
  }
  return B2G_OK;
}

// ---- the same pairs without allocating ------------------------------------------------------
// b2g_collide_pairs_noalloc stages the pairs, both transforms and the manifold records of at most
// B2G_NOALLOC_MAX_PAIRS pairs in one file-scope device array instead of per-call allocations.
// Each device holds its own copy of the array for the life of the process. In host code the
// array's name is not a device address: the address of each device's copy is looked up once
// (noalloc_scratch_init, at world creation, with that device current) and only that pointer is
// passed to kernels and copies. Every kernel writes each record it owns in full, so nothing has
// to be cleared between calls; calls that overlap on one device share the array and race.
constexpr int kNoallocPairs = B2G_NOALLOC_MAX_PAIRS;
constexpr size_t kNoallocXaOff = sizeof(int2) * kNoallocPairs;
constexpr size_t kNoallocXbOff = kNoallocXaOff + sizeof(float4) * kNoallocPairs;
constexpr size_t kNoallocOutOff = kNoallocXbOff + sizeof(float4) * kNoallocPairs;
constexpr size_t kNoallocBytes = kNoallocOutOff + sizeof(b2g_manifold) * kNoallocPairs;
static_assert(kNoallocXaOff % 256 == 0 && kNoallocXbOff % 256 == 0 && kNoallocOutOff % 256 == 0,
              "scratch regions start on 256-byte boundaries");

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

int noalloc_scratch_init(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) {
    set_error("device " + std::to_string(device) + " is outside the scratch table");
    return B2G_ERR_DEVICE;
  }
  if (g_scratch_ptr[device]) return B2G_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error("hipGetSymbolAddress(scratch) failed on device " + std::to_string(device) + ": " +
              hipGetErrorString(e));
    return B2G_ERR_DEVICE;
  }
  g_scratch_ptr[device] = p;
  return B2G_OK;
}

// manifold records in the C-ABI layout, so one copy returns them
__global__ void __launch_bounds__(kBlock) k_collide_given_records(int n, const int2* __restrict__ pairs,
                                                                  const float4* __restrict__ xa4,
                                                                  const float4* __restrict__ xb4,
                                                                  const ShapeV* __restrict__ shapes,
                                                                  b2g_manifold* __restrict__ out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int2 pr = pairs[k];
  Manifold m;
  evaluate(&m, shapes[pr.x], xf_of(xa4[k]), shapes[pr.y], xf_of(xb4[k]));
  b2g_manifold r;
  r.type = m.type;
  r.point_count = m.count;
  r.local_normal[0] = m.local_normal.x; r.local_normal[1] = m.local_normal.y;
  r.local_point[0] = m.local_point.x; r.local_point[1] = m.local_point.y;
  for (int p = 0; p < 2; ++p) {
    r.points[p].local_point[0] = m.lp[p].x; r.points[p].local_point[1] = m.lp[p].y;
    r.points[p].normal_impulse = 0.0f; r.points[p].tangent_impulse = 0.0f;
    r.points[p].id = m.id[p]; r.points[p].pad = 0;
  }
  out[k] = r;
}

int collide_fixture_pairs_noalloc(World* w, int n, const int32_t* pairs, const float* xfa, const float* xfb,
                                  b2g_manifold* out) {
  if (n > kNoallocPairs) {
    set_error("b2g_collide_pairs_noalloc: " + std::to_string(n) + " pairs exceed the scratch area's " +
              std::to_string(kNoallocPairs));
    return B2G_ERR_CAPACITY;
  }
  if (n <= 0) return B2G_OK;
  int st = check_given_pairs(w, n, pairs, "b2g_collide_pairs_noalloc");
  if (st) return st;
  void* base = (w->device >= 0 && w->device < kMaxDevices) ? g_scratch_ptr[w->device] : nullptr;
  if (!base) {
    set_error("b2g_collide_pairs_noalloc: no scratch address for device " + std::to_string(w->device));
    return B2G_ERR_DEVICE;
  }
  unsigned char* b = static_cast<unsigned char*>(base);
  int2* dp = reinterpret_cast<int2*>(b);
  float4* da = reinterpret_cast<float4*>(b + kNoallocXaOff);
  float4* db = reinterpret_cast<float4*>(b + kNoallocXbOff);
  b2g_manifold* dout = reinterpret_cast<b2g_manifold*>(b + kNoallocOutOff);
  bool ok = hipMemcpyAsync(dp, pairs, sizeof(int2) * n, hipMemcpyHostToDevice, w->stream) == hipSuccess &&
            hipMemcpyAsync(da, xfa, sizeof(float4) * n, hipMemcpyHostToDevice, w->stream) == hipSuccess &&
            hipMemcpyAsync(db, xfb, sizeof(float4) * n, hipMemcpyHostToDevice, w->stream) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(k_collide_given_records, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, w->stream, n, dp, da,
                       db, w->f.shapev.p, dout);
    ok = hipGetLastError() == hipSuccess &&
         hipMemcpyAsync(out, dout, sizeof(b2g_manifold) * n, hipMemcpyDeviceToHost, w->stream) == hipSuccess &&
         hipStreamSynchronize(w->stream) == hipSuccess;
  }
  if (!ok) {
    set_error("b2g_collide_pairs_noalloc: device error");
    return B2G_ERR_DEVICE;
  }
  return B2G_OK;
}

}  // namespace b2g
