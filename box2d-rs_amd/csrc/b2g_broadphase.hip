// b2g broadphase: B2body::synchronize_fixtures + B2broadPhase::update_pairs rebuilt as a
// Morton-code LBVH over the proxies' fat AABBs.
//
// Reference behaviour kept (box2d-rs 0.0.4):
//   * swept AABB and fat-AABB rule of B2fixture::synchronize / B2dynamicTree::move_proxy
//     (src/private/dynamics/b2_fixture.rs:147-173, src/private/collision/b2_dynamic_tree.rs:109-168)
//     are evaluated bit-exactly per proxy;
//   * the contact set the reference holds when the next collide runs equals the set of proxy
//     pairs whose fat AABBs overlap and that pass add_pair's filters
//     (src/private/dynamics/b2_contact_manager.rs:178-300): proxies start moved, fat AABBs only
//     change when a proxy is moved, and collide destroys pairs whose fat AABBs stopped
//     overlapping. That set does not depend on the tree, so the GPU rebuilds it each step from
//     an LBVH instead of maintaining the reference's incremental SAH tree.
//
// Each pair is emitted by one proxy (its emitter) into that proxy's segment, found by
// count -> exclusive scan -> write, so the pair list is grouped by emitter and deterministic;
// each new pair finds its predecessor in the previous step's segment of the same emitter
// (warm starting). Pairs are stored as (emitter, other).
#include "b2g_internal.h"

namespace b2g {

static constexpr int kBlock = 256;

// ---- fixture synchronisation --------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_synchronize(int np, const int* __restrict__ fbody,
                                                        const int4* __restrict__ fshape, const float2* __restrict__ pool,
                                                        const int* __restrict__ btype, const float4* __restrict__ pos0,
                                                        const float4* __restrict__ bmass, const float4* __restrict__ bxf,
                                                        float4* __restrict__ faabb, float4* __restrict__ ffat,
                                                        int* __restrict__ moved) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int b = fbody[i];
  if (btype[b] == 0) { moved[i] = 0; return; }
  ShapeV s = load_shape(fshape[i], pool);
  float4 p0 = pos0[b];
  float4 ms = bmass[b];
  Xf xf1;
  xf1.q = rot(p0.z);
  xf1.p = v2(p0.x, p0.y) - mul(xf1.q, v2(ms.z, ms.w));
  Xf xf2 = xf_of(bxf[b]);
  Aabb a1 = shape_aabb(s, xf1);
  Aabb a2 = shape_aabb(s, xf2);
  Aabb a = combine(a1, a2);
  faabb[i] = make_float4(a.lo.x, a.lo.y, a.hi.x, a.hi.y);
  V2 disp = center(a2) - center(a1);
  // B2dynamicTree::move_proxy
  const V2 r = v2(kAabbExtension, kAabbExtension);
  Aabb fat;
  fat.lo = a.lo - r;
  fat.hi = a.hi + r;
  V2 d = kAabbMultiplier * disp;
  if (d.x < 0.0f) fat.lo.x += d.x; else fat.hi.x += d.x;
  if (d.y < 0.0f) fat.lo.y += d.y; else fat.hi.y += d.y;
  float4 t = ffat[i];
  Aabb tree;
  tree.lo = v2(t.x, t.y);
  tree.hi = v2(t.z, t.w);
  if (contains(tree, a)) {
    Aabb huge;
    huge.lo = fat.lo - 4.0f * r;
    huge.hi = fat.hi + 4.0f * r;
    if (contains(huge, tree)) { moved[i] = 0; return; }
  }
  ffat[i] = make_float4(fat.lo.x, fat.lo.y, fat.hi.x, fat.hi.y);
  moved[i] = 1;
}

// ---- scene bounds and Morton codes ---------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_bounds(int n, const float4* __restrict__ fat, int* __restrict__ out) {
  float lx = kMaxFloat, ly = kMaxFloat, hx = -kMaxFloat, hy = -kMaxFloat;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 a = fat[i];
    float cx = 0.5f * (a.x + a.z), cy = 0.5f * (a.y + a.w);
    lx = fminf(lx, cx); ly = fminf(ly, cy); hx = fmaxf(hx, cx); hy = fmaxf(hy, cy);
  }
  for (int d = 32; d >= 1; d >>= 1) {
    lx = fminf(lx, __shfl_xor(lx, d, 64));
    ly = fminf(ly, __shfl_xor(ly, d, 64));
    hx = fmaxf(hx, __shfl_xor(hx, d, 64));
    hy = fmaxf(hy, __shfl_xor(hy, d, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&out[0], float_flip(lx));
    atomicMin(&out[1], float_flip(ly));
    atomicMax(&out[2], float_flip(hx));
    atomicMax(&out[3], float_flip(hy));
  }
}

__global__ void k_bounds_init(int* out) {
  out[0] = float_flip(kMaxFloat);
  out[1] = float_flip(kMaxFloat);
  out[2] = float_flip(-kMaxFloat);
  out[3] = float_flip(-kMaxFloat);
}

__device__ inline unsigned spread15(unsigned x) {
  x &= 0x7fffu;
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}

// key = world id in the high bits, then a Morton code of `mbits` bits per axis of the fat-AABB
// centre; sorting by it groups each world's proxies and orders them along a Z curve
__global__ void __launch_bounds__(kBlock) k_morton(int n, const float4* __restrict__ fat, const int* __restrict__ bounds,
                                                   const int* __restrict__ world, int mbits,
                                                   unsigned* __restrict__ keys, unsigned* __restrict__ vals) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float cells = float((1u << mbits) - 1u);
  float lx = float_unflip(bounds[0]), ly = float_unflip(bounds[1]);
  float hx = float_unflip(bounds[2]), hy = float_unflip(bounds[3]);
  float sx = hx > lx ? cells / (hx - lx) : 0.0f;
  float sy = hy > ly ? cells / (hy - ly) : 0.0f;
  float4 a = fat[i];
  float cx = 0.5f * (a.x + a.z), cy = 0.5f * (a.y + a.w);
  unsigned ux = unsigned(fminf(fmaxf((cx - lx) * sx, 0.0f), cells));
  unsigned uy = unsigned(fminf(fmaxf((cy - ly) * sy, 0.0f), cells));
  unsigned code = spread15(ux) | (spread15(uy) << 1);
  keys[i] = (unsigned(world[i]) << (2 * mbits)) | code;
  vals[i] = unsigned(i);
}

// ---- LBVH (Karras 2012) --------------------------------------------------------------------
// Nodes: internal 0..n-2, leaves n-1..2n-2 (leaf of sorted position k is n-1+k).
__device__ inline int lbvh_delta(const unsigned* __restrict__ keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  unsigned a = keys[i], b = keys[j];
  if (a == b) return 32 + __clz(unsigned(i) ^ unsigned(j));
  return __clz(a ^ b);
}

__global__ void __launch_bounds__(kBlock) k_lbvh_build(int n, const unsigned* __restrict__ keys, int wshift,
                                                       int2* __restrict__ child, int2* __restrict__ worlds,
                                                       int* __restrict__ parent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = (lbvh_delta(keys, n, i, i + 1) - lbvh_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = lbvh_delta(keys, n, i, i - d);
  int lmax = 2;
  while (lbvh_delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (lbvh_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = lbvh_delta(keys, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (lbvh_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int gamma = i + s * d + (d < 0 ? -1 : 0);
  int lo = i < j ? i : j, hi = i < j ? j : i;
  int left = (lo == gamma) ? (n - 1 + gamma) : gamma;
  int right = (hi == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
  child[i] = make_int2(left, right);
  worlds[i] = make_int2(int(keys[lo] >> wshift), int(keys[hi] >> wshift));
  parent[left] = i;
  parent[right] = i;
}

// bottom-up refit: the second child to arrive at a node computes its box and climbs on
__global__ void __launch_bounds__(kBlock) k_lbvh_refit(int n, const unsigned* __restrict__ vals,
                                                       const float4* __restrict__ fat, const int2* __restrict__ child,
                                                       const int* __restrict__ parent, float4* __restrict__ box,
                                                       int* __restrict__ flag) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int node = n - 1 + k;
  box[node] = fat[vals[k]];
  if (n == 1) return;
  int p = parent[node];
  while (p >= 0) {
    __threadfence();
    int prev = __hip_atomic_fetch_add(&flag[p], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == 0) return;  // sibling not done yet; it will continue
    int2 c = child[p];
    float4 a = box[c.x];
    float4 b = box[c.y];
    a = make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w));
    box[p] = a;
    p = p == 0 ? -1 : parent[p];
  }
}

// ---- pair search --------------------------------------------------------------------------
struct PairFilter {
  const int* __restrict__ fbody;
  const int* __restrict__ btype;
  const int2* __restrict__ filter;
  const int4* __restrict__ fshape;
  const int* __restrict__ world;
  __device__ bool ok(int q, int p) const {
    int bq = fbody[q], bp = fbody[p];
    if (bq == bp) return false;
    if (world[q] != world[p]) return false;  // batched worlds never interact
    // B2body::should_collide: at least one body dynamic (src/private/dynamics/b2_body.rs:391-416)
    if (btype[bq] != 2 && btype[bp] != 2) return false;
    // default contact filter (src/private/dynamics/b2_world_callbacks.rs:6-18)
    int2 fq = filter[q], fp = filter[p];
    if (fq.y == fp.y && fq.y != 0) {
      if (fq.y < 0) return false;
    } else {
      unsigned catq = fq.x & 0xffff, maskq = (unsigned(fq.x) >> 16) & 0xffff;
      unsigned catp = fp.x & 0xffff, maskp = (unsigned(fp.x) >> 16) & 0xffff;
      if ((maskq & catp) == 0 || (catq & maskp) == 0) return false;
    }
    // a contact class must exist for the shape pair
    return registry_swap(fshape[q].x & 0xff, fshape[p].x & 0xff) >= 0;
  }
};

static constexpr int kTravBlock = 64;
static constexpr int kStack = 64;

template <bool WRITE>
__global__ void __launch_bounds__(kTravBlock) k_traverse(int n, const unsigned* __restrict__ vals,
                                                         const float4* __restrict__ fat, const int2* __restrict__ child,
                                                         const int2* __restrict__ worlds, int batched,
                                                         const float4* __restrict__ box, PairFilter pf,
                                                         int* __restrict__ count, const int* __restrict__ off,
                                                         int2* __restrict__ pairs, int* __restrict__ overflow) {
  __shared__ int stack[kStack][kTravBlock];
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int q = int(vals[k]);
  // Emitter rule: a pair is emitted once, by its non-static proxy; a pair of two non-static
  // proxies by the lower index. Static proxies (e.g. long container edges) emit nothing, so
  // no thread walks the hundreds of pairs a wall collects.
  const bool q_static = pf.btype[pf.fbody[q]] == 0;
  if (n == 1 || q_static) {
    if (!WRITE) count[q] = 0;
    return;
  }
  const float4 qa = fat[q];
  Aabb qb;
  qb.lo = v2(qa.x, qa.y);
  qb.hi = v2(qa.z, qa.w);
  int c = 0;
  int out = WRITE ? off[q] : 0;
  const int qw = pf.world[q];
  int sp = 0;
  stack[sp++][threadIdx.x] = 0;
  while (sp > 0) {
    int node = stack[--sp][threadIdx.x];
    float4 nb = box[node];
    Aabb b;
    b.lo = v2(nb.x, nb.y);
    b.hi = v2(nb.z, nb.w);
    if (!overlap(b, qb)) continue;
    if (node >= n - 1) {
      int p = int(vals[node - (n - 1)]);
      bool emit = p != q && (p > q || pf.btype[pf.fbody[p]] == 0);
      if (emit && pf.ok(q, p)) {
        if (WRITE) pairs[out + c] = make_int2(q, p);
        ++c;
      }
    } else {
      if (batched) {
        int2 wr = worlds[node];
        if (qw < wr.x || qw > wr.y) continue;
      }
      if (sp + 2 > kStack) { atomicExch(overflow, 1); break; }
      int2 ch = child[node];
      stack[sp++][threadIdx.x] = ch.y;
      stack[sp++][threadIdx.x] = ch.x;
    }
  }
  if (!WRITE) count[q] = c;
}

// predecessor of each new pair in the previous contact set (same emitter segment)
__global__ void __launch_bounds__(kBlock) k_merge(int n, const int2* __restrict__ pairs, int old_np,
                                                  const int* __restrict__ old_off, const int2* __restrict__ old_pairs,
                                                  int* __restrict__ old_index) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int2 pr = pairs[k];
  int found = -1;
  if (pr.x < old_np) {
    int s = old_off[pr.x], e = old_off[pr.x + 1];
    for (int i = s; i < e; ++i)
      if (old_pairs[i].y == pr.y) { found = i; break; }
  }
  old_index[k] = found;
}

// Pair search over the proxies' fat AABBs into contact set `dst` (pairs, offsets); returns the
// pair count in *npairs. Does not touch manifolds.
static int find_pairs(World* w, const float4* fat, int np, ContactArrays& dst, int* npairs) {
  BroadphaseArrays& bp = w->bp;
  int st;
  if ((st = bp.keys.reserve(np)) || (st = bp.vals.reserve(np)) || (st = bp.node_parent.reserve(2 * np)) ||
      (st = bp.node_child.reserve(np)) || (st = bp.node_worlds.reserve(np)) || (st = bp.node_aabb.reserve(2 * np)) ||
      (st = bp.node_flag.reserve(np)) ||
      (st = bp.count.reserve(np + 1)) || (st = dst.off.reserve(np + 1)) || (st = bp.bounds.reserve(2)))
    return st;
  int* bnd = reinterpret_cast<int*>(bp.bounds.p);
  int* ovf = bnd + 4;
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(1), 0, w->stream, bnd);
  B2G_HIP(hipMemsetAsync(ovf, 0, sizeof(int), w->stream));
  int gb = std::min(blocks_for(np, kBlock), 1024);
  hipLaunchKernelGGL(k_bounds, dim3(gb), dim3(kBlock), 0, w->stream, np, fat, bnd);
  // world bits above Morton bits: 30 Morton bits for one world, fewer for many
  int wbits = 0;
  while ((1 << wbits) < w->n_worlds) ++wbits;
  int mbits = std::min(15, (32 - wbits) / 2);
  int key_bits = wbits + 2 * mbits;
  hipLaunchKernelGGL(k_morton, dim3(blocks_for(np, kBlock)), dim3(kBlock), 0, w->stream, np, fat, bnd, w->f.world.p,
                     mbits, bp.keys.p, bp.vals.p);
  if ((st = radix_sort_pairs(w, bp.keys.p, bp.vals.p, np, key_bits))) return st;
  if (np > 1) {
    B2G_HIP(hipMemsetAsync(bp.node_parent.p, 0xff, sizeof(int) * (2 * np - 1), w->stream));
    hipLaunchKernelGGL(k_lbvh_build, dim3(blocks_for(np - 1, kBlock)), dim3(kBlock), 0, w->stream, np, bp.keys.p,
                       2 * mbits, bp.node_child.p, bp.node_worlds.p, bp.node_parent.p);
    B2G_HIP(hipMemsetAsync(bp.node_flag.p, 0, sizeof(int) * np, w->stream));
  }
  hipLaunchKernelGGL(k_lbvh_refit, dim3(blocks_for(np, kBlock)), dim3(kBlock), 0, w->stream, np, bp.vals.p, fat,
                     bp.node_child.p, bp.node_parent.p, bp.node_aabb.p, bp.node_flag.p);
  PairFilter pf{w->f.body.p, w->b.type.p, w->f.filter.p, w->f.shape.p, w->f.world.p};
  const int batched = wbits > 0;
  hipLaunchKernelGGL(k_traverse<false>, dim3(blocks_for(np, kTravBlock)), dim3(kTravBlock), 0, w->stream, np,
                     bp.vals.p, fat, bp.node_child.p, bp.node_worlds.p, batched, bp.node_aabb.p, pf, bp.count.p,
                     nullptr, nullptr, ovf);
  if ((st = exclusive_scan(w, bp.count.p, dst.off.p, np, dst.off.p + np))) return st;
  int h[2];
  B2G_HIP(hipMemcpyAsync(&h[0], dst.off.p + np, sizeof(int), hipMemcpyDeviceToHost, w->stream));
  B2G_HIP(hipMemcpyAsync(&h[1], ovf, sizeof(int), hipMemcpyDeviceToHost, w->stream));
  B2G_HIP(hipStreamSynchronize(w->stream));
  if (h[1]) {
    set_error("broadphase: LBVH traversal stack overflow");
    return B2G_ERR_CAPACITY;
  }
  int total = h[0];
  if ((st = dst.pair.reserve(std::max(total, 1)))) return st;
  hipLaunchKernelGGL(k_traverse<true>, dim3(blocks_for(np, kTravBlock)), dim3(kTravBlock), 0, w->stream, np,
                     bp.vals.p, fat, bp.node_child.p, bp.node_worlds.p, batched, bp.node_aabb.p, pf, nullptr,
                     dst.off.p, dst.pair.p, ovf);
  B2G_HIP(hipGetLastError());
  *npairs = total;
  return B2G_OK;
}

static int reserve_contacts(ContactArrays& c, int n) {
  int st;
  n = std::max(n, 1);
  if ((st = c.pair.reserve(n)) || (st = c.old_index.reserve(n)) || (st = c.ab.reserve(n)) || (st = c.hdr.reserve(n)) ||
      (st = c.geo.reserve(n)) || (st = c.pt0.reserve(n)) || (st = c.pt1.reserve(n)) || (st = c.ids.reserve(n)))
    return st;
  return B2G_OK;
}

int phase_synchronize(World* w) {
  if (w->np == 0) return B2G_OK;
  hipLaunchKernelGGL(k_synchronize, dim3(blocks_for(w->np, kBlock)), dim3(kBlock), 0, w->stream, w->np, w->f.body.p,
                     w->f.shape.p, w->f.pool.p, w->b.type.p, w->b.pos0.p, w->b.mass.p, w->b.xf.p, w->f.aabb.p,
                     w->f.fat.p, w->f.moved.p);
  B2G_HIP(hipGetLastError());
  return B2G_OK;
}

// New contact set from the current fat AABBs; becomes the current set, the old one is kept
// (its manifolds feed warm starting).
int phase_broadphase(World* w) {
  int next = 1 - w->cur;
  ContactArrays& dst = w->c[next];
  ContactArrays& old = w->c[w->cur];
  int n = 0;
  int st;
  if (w->np == 0) {
    w->cur = next;
    w->n_pairs = 0;
    return B2G_OK;
  }
  if ((st = find_pairs(w, w->f.fat.p, w->np, dst, &n))) return st;
  if ((st = reserve_contacts(dst, n))) return st;
  if (n > 0) {
    int old_np = w->set_np[w->cur];  // proxies when the old set was built (0: no old set)
    hipLaunchKernelGGL(k_merge, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, w->stream, n, dst.pair.p, old_np,
                       old.off.p, old.pair.p, dst.old_index.p);
    B2G_HIP(hipGetLastError());
  }
  w->set_np[next] = w->np;
  w->cur = next;
  w->n_pairs = n;
  return B2G_OK;
}

int broadphase_pairs_for(World* w, const float* fat_host, int64_t cap, int32_t* out, int64_t* count) {
  int np = w->np;
  DevArray<float4> fat;
  int st;
  if ((st = fat.reserve(std::max(np, 1)))) return st;
  B2G_HIP(hipMemcpyAsync(fat.p, fat_host, sizeof(float4) * np, hipMemcpyHostToDevice, w->stream));
  ContactArrays tmp;
  int n = 0;
  st = find_pairs(w, fat.p, np, tmp, &n);
  if (st == B2G_OK) {
    *count = n;
    int64_t m = std::min<int64_t>(cap, n);
    if (m > 0) {
      if (hipMemcpyAsync(out, tmp.pair.p, sizeof(int2) * m, hipMemcpyDeviceToHost, w->stream) != hipSuccess ||
          hipStreamSynchronize(w->stream) != hipSuccess) {
        set_error("broadphase pair readback failed");
        st = B2G_ERR_DEVICE;
      } else {
        for (int64_t i = 0; i < m; ++i)  // report as (lower, higher) fixture index
          if (out[2 * i] > out[2 * i + 1]) std::swap(out[2 * i], out[2 * i + 1]);
      }
    }
  }
  tmp.pair.release();
  tmp.off.release();
  fat.release();
  return st;
}

}  // namespace b2g
