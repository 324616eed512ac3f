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
// Each pair is emitted by one proxy (its emitter: the one with the lower sorted leaf position)
// into that proxy's segment, found by count -> exclusive scan -> write, so the pair list is
// grouped by emitter and deterministic; each new pair finds its predecessor in the previous
// step's segment of either of its proxies (the emitter of a pair changes when the two proxies
// swap places in the Morton order). Pairs are stored as (emitter, other).
#include "b2g_internal.h"

namespace b2g {

static constexpr int kBlock = 256;

// ---- fixture synchronisation --------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_synchronize(int np, const int* __restrict__ fbody,
                                                        const ShapeV* __restrict__ shapes,
                                                        const int* __restrict__ btype, const float4* __restrict__ pos0,
                                                        const float4* __restrict__ bmass, const float4* __restrict__ bxf,
                                                        float4* __restrict__ faabb, float4* __restrict__ ffat,
                                                        int* __restrict__ moved, IslandView iv,
                                                        int* __restrict__ any_moved) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int b = fbody[i];
  // only bodies of this step's islands are synchronised (src/private/dynamics/b2_world.rs:482-500)
  if (btype[b] == 0 || iv.asleep(b)) { moved[i] = 0; return; }
  if (!sync_proxy(i, b, shapes[i], pos0, bmass, bxf, faabb, ffat)) { moved[i] = 0; return; }
  moved[i] = 1;
  if (any_moved) *any_moved = 1;  // every writer stores 1
}

// ---- scene bounds and Morton codes ---------------------------------------------------------
// The scene's extent for the Morton codes, in two steps without atomics: each block of k_bounds
// writes the box of its proxies' fat-AABB centres to partial[block], and every block of k_morton
// reduces those partial boxes itself before coding its proxies (min and max are exact, so the
// result does not depend on the order). Block 0 also clears the pair search's flags.
static constexpr int kBoundsBlocks = 256;  // most partial boxes (one per k_bounds block)
__device__ inline void pair_flags_init(int* flags, const int* scan_ctl) {
  flags[0] = 0;                           // traversal stack overflow
  flags[1] = 0;                           // proxies whose pairs overflow their slots
  flags[2] = scan_ctl ? scan_ctl[1] : 0;  // a single-pass scan gave up waiting (read back with the flags)
}
__device__ inline void block_box(float& lx, float& ly, float& hx, float& hy) {
  __shared__ float red[4][kBlock / 64];
  for (int d = 32; d >= 1; d >>= 1) {
    lx = fminf(lx, __shfl_xor(lx, d, 64));
    ly = fminf(ly, __shfl_xor(ly, d, 64));
    hx = fmaxf(hx, __shfl_xor(hx, d, 64));
    hy = fmaxf(hy, __shfl_xor(hy, d, 64));
  }
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { red[0][wid] = lx; red[1][wid] = ly; red[2][wid] = hx; red[3][wid] = hy; }
  __syncthreads();
  for (int w = 0; w < kBlock / 64; ++w) {
    lx = fminf(lx, red[0][w]); ly = fminf(ly, red[1][w]); hx = fmaxf(hx, red[2][w]); hy = fmaxf(hy, red[3][w]);
  }
}

__global__ void __launch_bounds__(kBlock) k_bounds(int n, const int* __restrict__ list, const float4* __restrict__ fat,
                                                   float4* __restrict__ partial, int* __restrict__ flags,
                                                   const int* __restrict__ scan_ctl) {
  float lx = kMaxFloat, ly = kMaxFloat, hx = -kMaxFloat, hy = -kMaxFloat;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 a = fat[list[i]];
    float cx = 0.5f * (a.x + a.z), cy = 0.5f * (a.y + a.w);
    lx = fminf(lx, cx); ly = fminf(ly, cy); hx = fmaxf(hx, cx); hy = fmaxf(hy, cy);
  }
  block_box(lx, ly, hx, hy);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_float4(lx, ly, hx, hy);
  if (blockIdx.x == 0 && threadIdx.x == 0) pair_flags_init(flags, scan_ctl);
}

// the pair search's flags for a search without tree proxies (no k_bounds launch)
__global__ void k_pair_flags_init(int* flags, const int* __restrict__ scan_ctl) { pair_flags_init(flags, scan_ctl); }

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
__global__ void __launch_bounds__(kBlock) k_morton(int n, const int* __restrict__ list, const float4* __restrict__ fat,
                                                   const float4* __restrict__ partial, int nparts,
                                                   const int* __restrict__ world, int mbits, unsigned* __restrict__ keys,
                                                   unsigned* __restrict__ vals) {
  float lx = kMaxFloat, ly = kMaxFloat, hx = -kMaxFloat, hy = -kMaxFloat;
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) {
    const float4 b = partial[p];
    lx = fminf(lx, b.x); ly = fminf(ly, b.y); hx = fmaxf(hx, b.z); hy = fmaxf(hy, b.w);
  }
  block_box(lx, ly, hx, hy);
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = list[k];
  const float cells = float((1u << mbits) - 1u);
  float sx = hx > lx ? cells / (hx - lx) : 0.0f;
  float sy = hy > ly ? cells / (hy - ly) : 0.0f;
  float4 a = fat[i];
  float cx = 0.5f * (a.x + a.z), cy = 0.5f * (a.y + a.w);
  unsigned ux = unsigned(fminf(fmaxf((cx - lx) * sx, 0.0f), cells));
  unsigned uy = unsigned(fminf(fmaxf((cy - ly) * sy, 0.0f), cells));
  unsigned code = spread15(ux) | (spread15(uy) << 1);
  // indexed by list position: the sort runs over the n listed proxies
  keys[k] = (unsigned(world[i]) << (2 * mbits)) | code;
  vals[k] = unsigned(i);
}

// ---- BVH -------------------------------------------------------------------------------------
// Topology: Karras's binary radix tree over the sorted keys (internal node i splits at the
// highest differing key bit, so a node covers a compact Z-order cell; internal 0..n-2, the
// leaf of sorted position k is referenced as ~k). Boxes: each child covers a contiguous range
// of sorted leaves, and its box is the union over that range, read from a segment tree built
// level by level in LDS (level 0 = the sorted leaf boxes; node (L, i) = (L-1, 2i) u
// (L-1, 2i+1)) and stored in its parent's record. Neither step needs atomics or communication
// between workgroups, so the build is a handful of short launches and the tree is a pure
// function of the sorted order.
struct Levels {
  int count;
  int size[32];
  int off[32];
};

static constexpr int kReduceLevels = 8;

// What the pair filters need of a proxy, packed for one 16-byte load per visited leaf:
// {proxy, body, category | mask << 16, (group & 0xffff) | shape type << 16 | body type << 24}.
struct LeafInputs {
  const int* fbody;
  const int* btype;
  const int2* filter;
  const int4* fshape;
  const int* nc_flag;  // per body: joined by a joint that forbids collision (null: no such joint)
};
// leaf record: proxy, body, filter bits, group | shape type << 16 | body type << 24 | 1 << 30
// when a joint keeps the body from colliding with some other body
__device__ inline int4 leaf_record(int p, const LeafInputs& li) {
  const int b = li.fbody[p];
  const int2 f = li.filter[p];
  const int nc = li.nc_flag ? li.nc_flag[b] : 0;
  return make_int4(p, b, f.x,
                   (f.y & 0xffff) | ((li.fshape[p].x & 0xff) << 16) | ((li.btype[b] & 0x3f) << 24) | (nc << 30));
}

__global__ void __launch_bounds__(kBlock) k_seg_reduce(Levels lv, int L0, const unsigned* __restrict__ vals,
                                                       const float4* __restrict__ fat, float4* __restrict__ seg,
                                                       LeafInputs li, int4* __restrict__ leafrec) {
  __shared__ float4 sb[kBlock];
  const int t = threadIdx.x;
  const int i = blockIdx.x * kBlock + t;
  float4 e = make_float4(kMaxFloat, kMaxFloat, -kMaxFloat, -kMaxFloat);
  if (i < lv.size[L0]) {
    if (L0 == 0) {
      const int p = int(vals[i]);
      e = fat[p];
      seg[i] = e;
      leafrec[i] = leaf_record(p, li);
    } else {
      e = seg[lv.off[L0] + i];
    }
  }
  sb[t] = e;
  __syncthreads();
  for (int j = 1; j <= kReduceLevels; ++j) {
    const int L = L0 + j;
    if (L >= lv.count) break;
    const int cnt = kBlock >> j;
    float4 r = make_float4(kMaxFloat, kMaxFloat, -kMaxFloat, -kMaxFloat);
    if (t < cnt) {
      float4 a = sb[2 * t], b = sb[2 * t + 1];
      r = make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w));
    }
    __syncthreads();
    if (t < cnt) {
      sb[t] = r;
      const int gi = blockIdx.x * cnt + t;
      if (gi < lv.size[L]) seg[lv.off[L] + gi] = r;
    }
    __syncthreads();
  }
}

__device__ inline int lbvh_delta(const unsigned* __restrict__ keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  unsigned a = keys[i], b = keys[j];
  if (a == b) return 32 + __clz(unsigned(i) ^ unsigned(j));
  return __clz(a ^ b);
}

__device__ inline void box_union(float4& r, float4 b) {
  r = make_float4(fminf(r.x, b.x), fminf(r.y, b.y), fmaxf(r.z, b.z), fmaxf(r.w, b.w));
}

// union of sorted leaves [a, b): bottom-up segment-tree query
__device__ inline float4 seg_range(const float4* __restrict__ seg, const Levels& lv, int a, int b) {
  float4 r = make_float4(kMaxFloat, kMaxFloat, -kMaxFloat, -kMaxFloat);
  for (int L = 0; a < b; ++L) {
    if (a & 1) box_union(r, seg[lv.off[L] + a++]);
    if (b & 1) box_union(r, seg[lv.off[L] + --b]);
    a >>= 1;
    b >>= 1;
  }
  return r;
}

// Karras 2012 internal node i: split and children, then each child's box (a leaf's own box,
// or a segment-tree query over the child's leaf range) and world range
__global__ void __launch_bounds__(kBlock) k_lbvh_build(int n, const unsigned* __restrict__ keys, int wshift, Levels lv,
                                                       const float4* __restrict__ seg, BvhNode* __restrict__ nodes) {
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
  const bool left_leaf = lo == gamma, right_leaf = hi == gamma + 1;
  BvhNode nd;
  nd.box[0] = left_leaf ? seg[gamma] : seg_range(seg, lv, lo, gamma + 1);
  nd.box[1] = right_leaf ? seg[gamma + 1] : seg_range(seg, lv, gamma + 1, hi + 1);
  nd.link = make_int4(left_leaf ? ~gamma : gamma, right_leaf ? ~(gamma + 1) : gamma + 1, gamma, hi);
  nd.worlds = make_int4(int(keys[lo] >> wshift), int(keys[gamma] >> wshift), int(keys[gamma + 1] >> wshift),
                        int(keys[hi] >> wshift));
  nodes[i] = nd;
}

// Internal node i as a 4-wide node: each binary child that is internal is replaced by its two
// children (their boxes and ranges come from the child's record); a leaf child stays as one slot.
__global__ void __launch_bounds__(kBlock) k_bvh4_collapse(int n_internal, const BvhNode* __restrict__ nodes,
                                                          Bvh4Node* __restrict__ nodes4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_internal) return;
  const BvhNode nd = nodes[i];
  Bvh4Node out;
  const int link2[2] = {nd.link.x, nd.link.y};
  const int last2[2] = {nd.link.z, nd.link.w};
  const int wlo2[2] = {nd.worlds.x, nd.worlds.z}, whi2[2] = {nd.worlds.y, nd.worlds.w};
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    if (link2[s] < 0) {
      out.box[2 * s] = nd.box[s];
      out.link[2 * s] = link2[s];
      out.last[2 * s] = last2[s];
      out.wlo[2 * s] = wlo2[s];
      out.whi[2 * s] = whi2[s];
      out.box[2 * s + 1] = make_float4(kMaxFloat, kMaxFloat, -kMaxFloat, -kMaxFloat);
      out.link[2 * s + 1] = 0;
      out.last[2 * s + 1] = -1;
      out.wlo[2 * s + 1] = 1;
      out.whi[2 * s + 1] = 0;
    } else {
      const BvhNode ch = nodes[link2[s]];
      out.box[2 * s] = ch.box[0];
      out.box[2 * s + 1] = ch.box[1];
      out.link[2 * s] = ch.link.x;
      out.link[2 * s + 1] = ch.link.y;
      out.last[2 * s] = ch.link.z;
      out.last[2 * s + 1] = ch.link.w;
      out.wlo[2 * s] = ch.worlds.x;
      out.whi[2 * s] = ch.worlds.y;
      out.wlo[2 * s + 1] = ch.worlds.z;
      out.whi[2 * s + 1] = ch.worlds.w;
    }
  }
  nodes4[i] = out;
}

// ---- pair search --------------------------------------------------------------------------
// B2body::should_collide's joint rule (src/private/dynamics/b2_body.rs:398-412): bodies joined by a
// joint with collide_connected off do not collide
__device__ inline bool joined_no_collide(int ba, int bb, const int* __restrict__ nc_off,
                                         const int* __restrict__ nc_list) {
  for (int i = nc_off[ba], e = nc_off[ba + 1]; i < e; ++i)
    if (nc_list[i] == bb) return true;
  return false;
}

struct PairFilter {
  const int* __restrict__ fbody;
  const int* __restrict__ btype;
  const int2* __restrict__ filter;
  const int4* __restrict__ fshape;
  const int* __restrict__ world;
  const int* __restrict__ nc_off;   // null: no joint forbids a collision
  const int* __restrict__ nc_list;
  __device__ bool ok(int q, int p) const {
    int bq = fbody[q], bp = fbody[p];
    if (bq == bp) return false;
    if (world[q] != world[p]) return false;  // batched worlds never interact
    // B2body::should_collide: at least one body dynamic (src/private/dynamics/b2_body.rs:391-416)
    if (btype[bq] != 2 && btype[bp] != 2) return false;
    if (nc_off && joined_no_collide(bq, bp, nc_off, nc_list)) return false;
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

static constexpr int kTravBlock = 64;                  // one wave: 16 queries of 4 lanes
static constexpr int kTravLanes = 4;                   // lanes per query, one per slot of a 4-wide node
static constexpr int kTravQueries = kTravBlock / kTravLanes;
static constexpr int kStack = 64;
static constexpr int kSlots = 64;

// PairFilter::ok on two leaf records of one world (the node world ranges have already matched
// the worlds): same body, at least one dynamic body, the default contact filter, and a contact
// class for the shape pair.
__device__ inline bool leaf_pair_ok(int4 a, int4 b, const int* __restrict__ nc_off, const int* __restrict__ nc_list) {
  if (a.y == b.y) return false;
  if (((a.w >> 24) & 0x3f) != 2 && ((b.w >> 24) & 0x3f) != 2) return false;
  if (((a.w & b.w) >> 30) & 1)
    if (joined_no_collide(a.y, b.y, nc_off, nc_list)) return false;
  const int ga = static_cast<short>(a.w & 0xffff), gb = static_cast<short>(b.w & 0xffff);
  if (ga == gb && ga != 0) {
    if (ga < 0) return false;
  } else {
    const unsigned cata = a.z & 0xffff, maska = (unsigned(a.z) >> 16) & 0xffff;
    const unsigned catb = b.z & 0xffff, maskb = (unsigned(b.z) >> 16) & 0xffff;
    if ((maska & catb) == 0 || (cata & maskb) == 0) return false;
  }
  return registry_swap((a.w >> 16) & 0xff, (b.w >> 16) & 0xff) >= 0;
}

struct TreeView {
  int n;                       // leaves in the tree
  const Bvh4Node* nodes;       // internal nodes as 4-wide nodes
  const int4* leafrec;         // sorted leaf records
  const int* loose_off;        // per world: range of big static proxies kept out of the tree
  const int* loose;
  int n_loose;                 // big static proxies over all worlds
  int batched;
};

// One tree walk per tree proxy, walked by 4 lanes: query k = sorted leaf k (neighbouring groups
// walk neighbouring paths), lane s of the group tests slot s of each visited node. The emitter of
// a pair of tree proxies is the one with the lower sorted position, so a walk skips every child
// whose leaves all sit at or before its own position: the last leaf under each slot is in the
// node record. A visited node is the 4-wide record of a binary node (its grandchildren's boxes),
// so the walk descends two levels per 128-byte record, read as four 32-byte slot loads by the
// group, and a leaf is tested from its parent's record without a load of its own. The group's
// stack lives in LDS; its depth, the pair count and the order of emitted pairs (slot order within
// a node) are the same in the four lanes, kept in step by ballots over the wave. Then the world's
// big static proxies (outside the tree) are tested directly, four at a time; every tree proxy
// emits its pairs with them. First pass: the first kSlots pairs go to the query's slots and the
// count is recorded. The overflow pass re-walks only queries with more than kSlots pairs and
// writes their whole segment; the walk order is deterministic, so both passes list the pairs in
// the same order.
template <bool OVERFLOW_PASS>
__global__ void __launch_bounds__(kTravBlock) k_traverse(int nq, const float4* __restrict__ fat, TreeView tv,
                                                         PairFilter pf, int* __restrict__ count,
                                                         int* __restrict__ slots, const int* __restrict__ off,
                                                         int2* __restrict__ pairs, int* __restrict__ flags) {
  __shared__ int stack[kStack][kTravQueries];
  const int k = (blockIdx.x * blockDim.x + threadIdx.x) / kTravLanes;
  const int g = threadIdx.x / kTravLanes;  // the query's stack column (and lane group of the wave)
  const int s = threadIdx.x % kTravLanes;  // the node slot this lane tests
  if (k >= nq) return;
  // proxies outside the tree emit nothing: their counts are cleared here
  if (!OVERFLOW_PASS && s == 0)
    for (int e = k; e < tv.n_loose; e += nq) count[tv.loose[e]] = 0;
  const int4 qr = tv.leafrec[k];
  const int q = qr.x;
  if (OVERFLOW_PASS && count[q] <= kSlots) return;
  const float4 qa = fat[q];
  Aabb qb;
  qb.lo = v2(qa.x, qa.y);
  qb.hi = v2(qa.z, qa.w);
  const int qw = pf.world[q];
  const unsigned below = (1u << s) - 1u;  // this lane's lower slots in the group mask
  int c = 0;                              // pairs emitted so far (same in the group's lanes)
  const int out = OVERFLOW_PASS ? off[q] : 0;
  // emit the group's hits in slot order
  auto emit = [&](bool hit, int p) {
    const unsigned gm = unsigned(__ballot(hit) >> (g * kTravLanes)) & 0xfu;
    if (hit) {
      const int r = c + __popc(gm & below);
      if (OVERFLOW_PASS) pairs[out + r] = make_int2(q, p);
      else if (r < kSlots) slots[size_t(r) * nq + k] = p;
    }
    c += __popc(gm);
  };
  if (tv.n > 1) {
    if (s == 0) stack[0][g] = 0;
    int sp = 1;
    while (sp > 0) {
      const Bvh4Node* nd = tv.nodes + stack[--sp][g];
      const float4 cb = nd->box[s];
      const int ch = nd->link[s];
      bool hit = nd->last[s] > k;
      if (tv.batched) hit = hit && qw >= nd->wlo[s] && qw <= nd->whi[s];
      Aabb b;
      b.lo = v2(cb.x, cb.y);
      b.hi = v2(cb.z, cb.w);
      hit = hit && overlap(b, qb);
      bool ok = false;
      int p = 0;
      if (hit && ch < 0) {
        const int4 pr = tv.leafrec[~ch];
        ok = leaf_pair_ok(qr, pr, pf.nc_off, pf.nc_list);
        p = pr.x;
      }
      emit(ok, p);
      // internal children: pushed so that the lowest slot is walked first
      const bool inner = hit && ch >= 0;
      const unsigned im = unsigned(__ballot(inner) >> (g * kTravLanes)) & 0xfu;
      const int n_in = __popc(im);
      if (sp + n_in > kStack) {
        if (s == 0) atomicExch(&flags[0], 1);
        break;
      }
      if (inner) stack[sp + n_in - 1 - __popc(im & below)][g] = ch;
      sp += n_in;
    }
  }
  for (int e0 = tv.loose_off[qw], end = tv.loose_off[qw + 1]; e0 < end; e0 += kTravLanes) {
    const int e = e0 + s;
    bool ok = false;
    int p = 0;
    if (e < end) {
      p = tv.loose[e];
      const float4 nb = fat[p];
      Aabb b;
      b.lo = v2(nb.x, nb.y);
      b.hi = v2(nb.z, nb.w);
      ok = overlap(b, qb) && pf.ok(q, p);
    }
    emit(ok, p);
  }
  if (!OVERFLOW_PASS && s == 0) {
    count[q] = c;
    if (c > kSlots) atomicAdd(&flags[1], 1);
  }
}

__global__ void __launch_bounds__(kBlock) k_pairs_from_slots(int nq, const int4* __restrict__ leafrec,
                                                             const int* __restrict__ count,
                                                             const int* __restrict__ slots, const int* __restrict__ off,
                                                             int2* __restrict__ pairs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nq) return;
  const int q = leafrec[k].x;
  const int c = min(count[q], kSlots);
  const int o = off[q];
  for (int j = 0; j < c; ++j) pairs[o + j] = make_int2(q, slots[size_t(j) * nq + k]);
}

// predecessor of each new pair in the previous contact set: the pair sits in the old segment of
// one of its two proxies (its emitter then)
__device__ inline int find_in_segment(int a, int b, int old_np, const int* __restrict__ old_off,
                                      const int2* __restrict__ old_pairs) {
  if (a >= old_np) return -1;
  for (int i = old_off[a], e = old_off[a + 1]; i < e; ++i)
    if (old_pairs[i].y == b) return i;
  return -1;
}
// (also clears the new set's solve colours: the solve of this set writes those of its touching
// contacts)
__global__ void __launch_bounds__(kBlock) k_merge(int n, const int2* __restrict__ pairs, int old_np,
                                                  const int* __restrict__ old_off, const int2* __restrict__ old_pairs,
                                                  int* __restrict__ old_index, int* __restrict__ colour) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  colour[k] = -1;
  int2 pr = pairs[k];
  int found = find_in_segment(pr.x, pr.y, old_np, old_off, old_pairs);
  if (found < 0) found = find_in_segment(pr.y, pr.x, old_np, old_off, old_pairs);
  old_index[k] = found;
}

// Pair search over the proxies' fat AABBs into contact set `dst` (pairs, offsets); returns the
// pair count in *npairs. Does not touch manifolds.
static int find_pairs(World* w, const float4* fat, int np, ContactArrays& dst, int* npairs) {
  BroadphaseArrays& bp = w->bp;
  const int nt = w->n_tree;  // proxies in the tree (the others are big statics, tested directly)
  int st;
  Levels lv;
  lv.count = 0;
  int total_nodes = 0;
  for (int sz = std::max(nt, 1);; sz = (sz + 1) / 2) {
    if (lv.count == 32) { set_error("broadphase: too many proxies"); return B2G_ERR_CAPACITY; }
    lv.size[lv.count] = sz;
    lv.off[lv.count] = total_nodes;
    total_nodes += sz;
    ++lv.count;
    if (sz <= 1) break;
  }
  if ((st = bp.keys.reserve(std::max(nt, 1))) || (st = bp.vals.reserve(std::max(nt, 1))) ||
      (st = bp.seg.reserve(total_nodes)) || (st = bp.nodes.reserve(std::max(nt, 1))) ||
      (st = bp.nodes4.reserve(std::max(nt, 1))) ||
      (st = bp.leafrec.reserve(std::max(nt, 1))) ||
      (st = bp.count.reserve(np + 1)) || (st = bp.slots.reserve(size_t(np) * kSlots)) ||
      (st = dst.off.reserve(np + 1)) || (st = bp.bounds.reserve(1 + kBoundsBlocks)))
    return st;
  // bounds[0]: the pair search's flags ([0] stack overflow, [1] proxies with more pairs than
  // slots, [2] scan error, [3] pair total); bounds[1..]: k_bounds' partial boxes
  int* flags = reinterpret_cast<int*>(bp.bounds.p);
  float4* partial = bp.bounds.p + 1;
  const int* scan_ctl = static_cast<const int*>(bp.scan_ctl.p);
  if (nt == 0) hipLaunchKernelGGL(k_pair_flags_init, dim3(1), dim3(1), 0, w->stream, flags, scan_ctl);
  // world bits above Morton bits: 12 bits per axis (4096 cells across the scene's extent, finer
  // than any body on the scenes here; equal keys keep proxy order), fewer for many worlds, so one
  // world sorts in three 8-bit passes
  int wbits = 0;
  while ((1 << wbits) < w->n_worlds) ++wbits;
  const int mbits = std::min(12, (32 - wbits) / 2);
  const int key_bits = wbits + 2 * mbits;
  const bool nc = w->n_nocollide > 0;
  const LeafInputs li{w->f.body.p, w->b.type.p, w->f.filter.p, w->f.shape.p, nc ? w->j.nc_flag.p : nullptr};
  if (nt > 0) {
    const int gb = std::min(blocks_for(nt, kBlock), kBoundsBlocks);
    hipLaunchKernelGGL(k_bounds, dim3(gb), dim3(kBlock), 0, w->stream, nt, w->f.tree_list.p, fat, partial, flags,
                       scan_ctl);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(nt, kBlock)), dim3(kBlock), 0, w->stream, nt, w->f.tree_list.p, fat,
                       partial, gb, w->f.world.p, mbits, bp.keys.p, bp.vals.p);
    if ((st = radix_sort_pairs(w, bp.keys.p, bp.vals.p, nt, key_bits))) return st;
    for (int L0 = 0; L0 == 0 || L0 < lv.count - 1; L0 += kReduceLevels)
      hipLaunchKernelGGL(k_seg_reduce, dim3(blocks_for(lv.size[L0], kBlock)), dim3(kBlock), 0, w->stream, lv, L0,
                         bp.vals.p, fat, bp.seg.p, li, bp.leafrec.p);
    if (nt > 1) {
      hipLaunchKernelGGL(k_lbvh_build, dim3(blocks_for(nt - 1, kBlock)), dim3(kBlock), 0, w->stream, nt, bp.keys.p,
                         2 * mbits, lv, bp.seg.p, bp.nodes.p);
      hipLaunchKernelGGL(k_bvh4_collapse, dim3(blocks_for(nt - 1, kBlock)), dim3(kBlock), 0, w->stream, nt - 1,
                         bp.nodes.p, bp.nodes4.p);
    }
  }
  PairFilter pf{w->f.body.p, w->b.type.p, w->f.filter.p, w->f.shape.p, w->f.world.p, nc ? w->j.nc_off.p : nullptr,
                nc ? w->j.nc_list.p : nullptr};
  TreeView tv{nt, bp.nodes4.p, bp.leafrec.p, w->f.loose_off.p, w->f.loose.p, np - nt, wbits > 0};
  // tree proxies walk in Morton order; proxies outside the tree are static and emit nothing
  // (the walk clears their counts; with no tree proxies there is no walk)
  if (nt == 0 && np > 0) B2G_HIP(hipMemsetAsync(bp.count.p, 0, sizeof(int) * np, w->stream));
  if (nt > 0) {
    launch_timed(w, B2G_KERNEL_TRAVERSE, nt, k_traverse<false>, dim3(blocks_for(nt * kTravLanes, kTravBlock)),
                 dim3(kTravBlock),
                 nt, fat, tv, pf, bp.count.p, bp.slots.p, nullptr, nullptr, flags);
  }
  // the pair total also lands beside the flags (flags[3]), so one read-back brings both
  if ((st = exclusive_scan(w, bp.count.p, dst.off.p, np, dst.off.p + np, flags + 3))) return st;
  // flags: [0] stack overflow, [1] slot overflow, [2] scan error, [3] pair total
  int* f = w->h_counters + kPinnedPairFlags;  // pinned
  B2G_HIP(hipMemcpyAsync(f, flags, 4 * sizeof(int), hipMemcpyDeviceToHost, w->stream));
  B2G_HIP(hipStreamSynchronize(w->stream));
  const int h[4] = {f[3], f[0], f[1], f[2]};
  if ((st = kt_harvest(w))) return st;
  if (h[1]) {
    set_error("broadphase: BVH traversal stack overflow");
    return B2G_ERR_CAPACITY;
  }
  if (h[3]) {
    set_error("device scan: a tile waited too long for its predecessor");
    return B2G_ERR_DEVICE;
  }
  int total = h[0];
  kt_aux(w, B2G_KERNEL_TRAVERSE, total);
  if ((st = dst.pair.reserve(pair_capacity(w, total)))) return st;
  if (nt > 0)
    hipLaunchKernelGGL(k_pairs_from_slots, dim3(blocks_for(nt, kBlock)), dim3(kBlock), 0, w->stream, nt,
                       bp.leafrec.p, bp.count.p, bp.slots.p, dst.off.p, dst.pair.p);
  if (h[2] > 0)
    hipLaunchKernelGGL(k_traverse<true>, dim3(blocks_for(nt * kTravLanes, kTravBlock)), dim3(kTravBlock), 0, w->stream, nt, fat,
                       tv, pf, bp.count.p, bp.slots.p, dst.off.p, dst.pair.p, flags);
  B2G_HIP(hipGetLastError());
  *npairs = total;
  return B2G_OK;
}

static int reserve_contacts(World* w, ContactArrays& c, int pairs) {
  int st;
  const size_t n = pair_capacity(w, pairs);
  if ((st = c.pair.reserve(n)) || (st = c.old_index.reserve(n)) || (st = c.ab.reserve(n)) || (st = c.hdr.reserve(n)) ||
      (st = c.geo.reserve(n)) || (st = c.pt0.reserve(n)) || (st = c.pt1.reserve(n)) || (st = c.ids.reserve(n)) ||
      (st = c.colour.reserve(n)))
    return st;
  return B2G_OK;
}

int phase_synchronize_view(World* w, IslandView iv, int* any_moved) {
  if (w->np == 0) return B2G_OK;
  hipLaunchKernelGGL(k_synchronize, dim3(blocks_for(w->np, kBlock)), dim3(kBlock), 0, w->stream, w->np, w->f.body.p,
                     w->f.shapev.p, w->b.type.p, w->b.pos0.p, w->b.mass.p, w->b.xf.p, w->f.aabb.p,
                     w->f.fat.p, w->f.moved.p, iv, any_moved);
  B2G_HIP(hipGetLastError());
  return B2G_OK;
}

int phase_synchronize(World* w) {
  return phase_synchronize_view(w, IslandView{w->s.parent.p, w->sleep_possible() ? w->s.isl_awake.p : nullptr},
                                nullptr);
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
    w->n_pairs_old = w->n_pairs;
    w->n_pairs = 0;
    return B2G_OK;
  }
  if ((st = find_pairs(w, w->f.fat.p, w->np, dst, &n))) return st;
  if ((st = reserve_contacts(w, dst, n))) return st;
  if (n > 0) {
    int old_np = w->set_np[w->cur];  // proxies when the old set was built (0: no old set)
    hipLaunchKernelGGL(k_merge, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, w->stream, n, dst.pair.p, old_np,
                       old.off.p, old.pair.p, dst.old_index.p, dst.colour.p);
    B2G_HIP(hipGetLastError());
  }
  w->set_np[next] = w->np;
  w->cur = next;
  w->n_pairs_old = w->n_pairs;
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
