// b2g solver: the island solve of B2world::solve rebuilt as graph-coloured Gauss-Seidel.
//
// Reference (box2d-rs 0.0.4): src/private/dynamics/b2_world.rs:356-507 (islands),
// src/private/dynamics/b2_island_private.rs:129-328 (integrate, solve, copy back) and
// src/private/dynamics/b2_contact_solver_private.rs:20-730 (contact solver).
//
// The touching contacts are coloured so that no two contacts of one colour share a dynamic
// body; a colour's contacts are then independent and are solved in parallel, colour after
// colour, which is sequential Gauss-Seidel in colour order (a CPU replay in that order gives
// the same numbers). Islands are the connected components of the contact graph through
// non-static bodies (union-find); they carry the reference's per-island early exit of the
// position iterations. Colouring is Jones-Plassmann-Luby with hash priorities and first-fit
// colours, so it is deterministic. Body state is SoA float4 in HBM.
#include <climits>
#include <cstring>

#include "b2g_internal.h"

namespace b2g {

static constexpr int kBlock = 256;
static constexpr int kMaxColors = 128;
// solver counter after the colour histogram and the last colouring round: constraints in islands
// too large for the island solve's bins
static constexpr int kBigIslandSlot = 4 + kMaxColors + 1;

// ---- integration --------------------------------------------------------------------------
// src/private/dynamics/b2_island_private.rs:135-170
// Islands with no awake body are not solved: their bodies keep their state. Bodies of awake
// islands are set awake (the island DFS of src/private/dynamics/b2_world.rs:437-447, which does
// not reset the sleep time).

struct IntegrateArgs {
  const int* btype;
  const float4* pos;
  float4* vel;
  float4* pos0;
  const float4* props;
  const float4* bmass;
  const float4* force;
  float gx, gy, h;
};
__device__ inline void integrate_velocity_one(int b, const IntegrateArgs& a, const IslandView& iv, int* awake) {
  int t = a.btype[b];
  if (t == 0 || iv.asleep(b)) return;
  if (awake) awake[b] = 1;
  a.pos0[b] = a.pos[b];
  if (t != 2) return;
  float4 v = a.vel[b];
  float4 pr = a.props[b];  // mass, gravity scale, linear damping, angular damping
  float4 ms = a.bmass[b];
  float4 f = a.force[b];
  const float h = a.h;
  float s = h * ms.x;
  float gm = pr.y * pr.x;
  v.x += s * (gm * a.gx + f.x);
  v.y += s * (gm * a.gy + f.y);
  v.z += h * ms.y * f.z;
  float dl = 1.0f / (1.0f + h * pr.z);
  float da = 1.0f / (1.0f + h * pr.w);
  v.x *= dl;
  v.y *= dl;
  v.z *= da;
  a.vel[b] = v;
}
__global__ void __launch_bounds__(kBlock) k_integrate_velocities(int nb, IntegrateArgs a, IslandView iv,
                                                                 int* __restrict__ awake) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) integrate_velocity_one(b, a, iv, awake);
}

// src/private/dynamics/b2_island_private.rs:221-252
__device__ inline void integrate_position_one(int b, const int* __restrict__ btype, float4* __restrict__ pos,
                                              float4* __restrict__ vel, float h, const IslandView& iv) {
  if (btype[b] == 0 || iv.asleep(b)) return;
  float4 p = pos[b], v = vel[b];
  V2 tr = h * v2(v.x, v.y);
  if (dot(tr, tr) > kMaxTranslation * kMaxTranslation) {
    float ratio = kMaxTranslation / len(tr);
    v.x *= ratio;
    v.y *= ratio;
  }
  float rotation = h * v.z;
  if (rotation * rotation > kMaxRotation * kMaxRotation) {
    float ratio = kMaxRotation / fabsf(rotation);
    v.z *= ratio;
  }
  p.x += h * v.x;
  p.y += h * v.y;
  p.z += h * v.z;
  pos[b] = p;
  vel[b] = v;
}
__global__ void __launch_bounds__(kBlock) k_integrate_positions(int nb, const int* __restrict__ btype,
                                                                float4* __restrict__ pos, float4* __restrict__ vel, float h,
                                                                IslandView iv) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) integrate_position_one(b, btype, pos, vel, h, iv);
}

// copy back + synchronize_transform (src/private/dynamics/b2_island_private.rs:276-285,
// src/b2_body.rs:974-977). Static bodies keep the transform computed at creation (their
// angle never changes, so the reference recomputes the same value).
__device__ inline void finalize_one(int b, const int* __restrict__ btype, const float4* __restrict__ pos,
                                    const float4* __restrict__ bmass, float4* __restrict__ xf, const IslandView& iv) {
  if (btype[b] == 0 || iv.asleep(b)) return;
  float4 p = pos[b], ms = bmass[b];
  Rot q = rot(p.z);
  V2 o = v2(p.x, p.y) - mul(q, v2(ms.z, ms.w));
  xf[b] = make_float4(o.x, o.y, q.s, q.c);
}
__global__ void __launch_bounds__(kBlock) k_finalize(int nb, const int* __restrict__ btype, const float4* __restrict__ pos,
                                                     const float4* __restrict__ bmass, float4* __restrict__ xf,
                                                     IslandView iv) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) finalize_one(b, btype, pos, bmass, xf, iv);
}

// Island sleep (src/private/dynamics/b2_island_private.rs:291-327), in two passes over the bodies
// of awake islands: each body's sleep time is reset when it may not sleep or moves faster than
// the tolerances, else advanced by h; an island whose bodies all reached kTimeToSleep (no body
// marks it restless) and whose position iterations converged goes to sleep.
__global__ void __launch_bounds__(kBlock) k_sleep_time(int nb, const int* __restrict__ btype,
                                                       const int* __restrict__ auto_sleep, const float4* __restrict__ vel,
                                                       float* __restrict__ sleep_time, IslandView iv,
                                                       int* __restrict__ restless, float h) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb || btype[b] == 0 || iv.asleep(b)) return;
  const float lin_tol_sqr = kLinearSleepTolerance * kLinearSleepTolerance;
  const float ang_tol_sqr = kAngularSleepTolerance * kAngularSleepTolerance;
  const float4 v = vel[b];
  float t;
  if (!auto_sleep[b] || v.z * v.z > ang_tol_sqr || v.x * v.x + v.y * v.y > lin_tol_sqr) {
    t = 0.0f;
  } else {
    t = sleep_time[b] + h;
  }
  sleep_time[b] = t;
  if (!(t >= kTimeToSleep)) restless[iv.parent[b]] = 1;
}
__global__ void __launch_bounds__(kBlock) k_sleep_apply(int nb, const int* __restrict__ btype, IslandView iv,
                                                        const int* __restrict__ restless, const int* __restrict__ done,
                                                        int* __restrict__ awake, float* __restrict__ sleep_time,
                                                        float4* __restrict__ vel, float4* __restrict__ force) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb || btype[b] == 0 || iv.asleep(b)) return;
  const int r = iv.parent[b];
  if (restless[r] || !done[r]) return;
  // B2body::set_awake(false)
  awake[b] = 0;
  sleep_time[b] = 0.0f;
  vel[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  force[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ---- touching contacts, islands, colouring -----------------------------------------------------
// per-step solver counters: [0] touching (written by the scan), [1] uncoloured, [2] colour
// overflow, [3] unused, then the colour histogram; and the colour scatter cursors
__device__ inline void counters_init(int t, int* c, int* __restrict__ cursor) {
  if (t == 0) {
    c[1] = 0;
    c[2] = 0;
    c[3] = 0;
    c[kBigIslandSlot] = 0;
  }
  for (int i = t; i <= kMaxColors; i += kBlock) c[4 + i] = 0;
  for (int i = t; i < kMaxColors; i += kBlock) cursor[i] = 0;
}

// One thread per pair of the contact set: a touching pair becomes constraint j = its rank among
// the touching pairs (the exclusive scan of the touching flags), with its bodies, its share of
// the bodies' degrees and the solve colour of its predecessor in the previous contact set
// (old_colour null: no reuse). Block 0 also initialises the per-step solver counters.
__global__ void __launch_bounds__(kBlock) k_touching_setup(int n, const int* __restrict__ flag,
                                                           const int* __restrict__ scan, int* __restrict__ list,
                                                           const int2* __restrict__ ab, const int* __restrict__ fbody,
                                                           const int* __restrict__ btype, int2* __restrict__ bodies,
                                                           int* __restrict__ deg, const int* __restrict__ old_index,
                                                           const int* __restrict__ old_colour,
                                                           int* __restrict__ prev_colour, int* counters,
                                                           int* __restrict__ cursor) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0) counters_init(threadIdx.x, counters, cursor);
  if (k >= n || !flag[k]) return;
  const int j = scan[k];
  list[j] = k;
  int2 f = ab[k];
  int ba = fbody[f.x], bb = fbody[f.y];
  bodies[j] = make_int2(ba, bb);
  if (btype[ba] == 2) atomicAdd(&deg[ba], 1);
  if (btype[bb] == 2) atomicAdd(&deg[bb], 1);
  int pc = -1;
  if (old_colour) {
    const int o = old_index[k];
    if (o >= 0) pc = old_colour[o];
  }
  prev_colour[j] = pc >= 0 && pc < kMaxColors ? pc : -1;
}

// Kernels over the touching contacts read their count from device memory (*ncp) so that the
// host does not wait for the compaction; they are launched over the pair count, an upper bound.
__global__ void __launch_bounds__(kBlock) k_fill_adjacency(const int* __restrict__ ncp, const int2* __restrict__ bodies,
                                                           const int* __restrict__ btype, const int* __restrict__ off,
                                                           int* __restrict__ cursor, int* __restrict__ adj) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= *ncp) return;
  int2 b = bodies[j];
  if (btype[b.x] == 2) adj[off[b.x] + atomicAdd(&cursor[b.x], 1)] = j;
  if (btype[b.y] == 2) adj[off[b.y] + atomicAdd(&cursor[b.y], 1)] = j;
}

// per body: its own island root, no constraints yet; when every island is awake (no sleeping
// bodies possible) also the velocity integration, which then does not wait for the islands
__global__ void __launch_bounds__(kBlock) k_solver_body_init(int nb, int* __restrict__ parent, int* __restrict__ deg,
                                                             int* __restrict__ cursor, int integrate, IntegrateArgs ia,
                                                             int* __restrict__ isl_count) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) {
    parent[b] = b;
    deg[b] = 0;
    cursor[b] = 0;
    if (isl_count) isl_count[b] = 0;
    if (integrate) integrate_velocity_one(b, ia, IslandView{parent, nullptr}, nullptr);
  }
  if (b == 0) deg[nb] = 0;
}

// Per-step solver counters: [0] touching, [1] uncoloured, [2] colour overflow, [3] unused
// islands, then the colour histogram (kMaxColors counts and the last colouring round); also
// clears the colour scatter cursors.
__global__ void __launch_bounds__(kBlock) k_solver_counters_init(int* c, int* __restrict__ cursor) {
  counters_init(threadIdx.x, c, cursor);
}

// Find for the union kernel, where other threads link roots with atomicCAS at the same time.
// Reads are plain (cached, and on a multi-XCD part possibly stale within the launch): a parent
// pointer only ever moves to a node of higher rank, so a stale parent is still an ancestor and
// the walk only climbs. A root found from stale reads fails the linking CAS, which returns the
// node's true parent; the walk continues from there, so each failed CAS climbs at least one true
// level and no coherent (device-scope) load is needed. parent[] is written only by that CAS.
__device__ inline int uf_read(int* parent, int x) {
  return __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);  // a plain cached load
}
__device__ inline int uf_find(int* parent, int x) {
  for (int p = uf_read(parent, x); p != x; p = uf_read(parent, x)) x = p;
  return x;
}
// Linking order: the root with the lower (hash, index) goes under the other. A bijective hash
// makes this a random linking order, which keeps the expected tree depth logarithmic (linking by
// plain index builds chains as long as a row of touching bodies when a wave links neighbours in
// parallel). The final root of a component is its member with the highest (hash, index), a
// function of the component alone, so island ids do not depend on the order of the links.
__device__ inline unsigned uf_rank(int x) {
  unsigned h = unsigned(x) * 0x9E3779B1u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return h;
}
__device__ inline bool uf_below(int a, int c) {
  const unsigned ha = uf_rank(a), hc = uf_rank(c);
  return ha < hc || (ha == hc && a < c);
}
__device__ inline void union_bodies(int x, int y, const int* __restrict__ btype, int* parent) {
  if (btype[x] == 0 || btype[y] == 0) return;
  int a = uf_find(parent, x), c = uf_find(parent, y);
  for (;;) {
    if (a == c) return;
    const bool a_low = uf_below(a, c);
    const int lo = a_low ? a : c, hi = a_low ? c : a;
    const int old = atomicCAS(&parent[lo], lo, hi);
    if (old == lo) return;
    // lo was linked meanwhile: climb from its true parent
    const int up = uf_find(parent, old);
    if (a_low) a = up;
    else c = up;
  }
}
__global__ void __launch_bounds__(kBlock) k_union(const int* __restrict__ ncp, const int2* __restrict__ bodies,
                                                  const int* __restrict__ btype, int* parent) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= *ncp) return;
  int2 b = bodies[j];
  union_bodies(b.x, b.y, btype, parent);
}
// joints join islands like touching contacts (src/private/dynamics/b2_world.rs:466-490)
__global__ void __launch_bounds__(kBlock) k_joint_union(int nj, const int2* __restrict__ jb,
                                                        const int* __restrict__ btype, int* parent) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nj) return;
  const int2 b = jb[j];
  union_bodies(b.x, b.y, btype, parent);
}
void launch_joint_union(World* w) {
  if (w->nj == 0) return;
  hipLaunchKernelGGL(k_joint_union, dim3(blocks_for(w->nj, kBlock)), dim3(kBlock), 0, w->stream, w->nj,
                     w->j.bodies.p, w->b.type.p, w->s.parent.p);
}

__global__ void __launch_bounds__(kBlock) k_compress(int nb, int* parent, int* __restrict__ violated,
                                                     int* __restrict__ done, int* __restrict__ isl_awake,
                                                     int* __restrict__ restless) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  // Read-only walk to the root; each thread writes only its own entry. (Halving writes here
  // could land on another body's entry after that body stored its root and leave a non-root.)
  int r = b;
  for (int p = parent[r]; p != r; p = parent[r]) r = p;
  parent[b] = r;
  violated[b] = 0;
  done[b] = 0;
  if (isl_awake) {
    isl_awake[b] = 0;
    restless[b] = 0;
  }
}

// An island is solved when one of its bodies is awake (the reference seeds its island search
// only at awake bodies: src/private/dynamics/b2_world.rs:389-400).
__global__ void __launch_bounds__(kBlock) k_island_awake(int nb, const int* __restrict__ btype,
                                                         const int* __restrict__ awake, const int* __restrict__ parent,
                                                         int* __restrict__ isl_awake) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb && btype[b] != 0 && awake[b]) isl_awake[parent[b]] = 1;
}

__device__ inline unsigned long long priority(int j) {
  unsigned x = unsigned(j) * 0x9E3779B1u;
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return (static_cast<unsigned long long>(x) << 32) | unsigned(j);
}

// One Jones-Plassmann-Luby round: an uncoloured constraint whose priority beats every
// uncoloured neighbour (a constraint sharing a dynamic body) takes the smallest colour no
// coloured neighbour has. Neighbours coloured in this same round count as uncoloured, so the
// decision depends only on the states at the start of the round (written by earlier launches):
// whether a load sees a neighbour's new state or its old one makes no difference, and the
// state loads and stores need no coherence beyond the launch boundary (workgroup-scope relaxed
// atomics: plain cached accesses that the compiler may not tear or invent).
// A neighbour o in state so: an uncoloured (or same-round) neighbour of higher priority makes
// j wait; a coloured one marks its colour used.
__device__ inline void color_look(int o, int so, int round, unsigned long long pj, bool& lose,
                                  unsigned long long& used0, unsigned long long& used1) {
  if (so == -1 || (so >> 8) == round) {
    if (priority(o) > pj) lose = true;
  } else {
    const int c = so & 0xff;
    if (c < 64) used0 |= 1ull << c;
    else used1 |= 1ull << (c - 64);
  }
}
// The first kColorUnroll adjacency entries of both bodies are loaded together and then all
// their states together (two rounds of independent loads instead of a chain of dependent ones
// per neighbour); longer lists finish in a loop. The outcome does not depend on the order in
// which neighbours are looked at.
static constexpr int kColorUnroll = 8;
__device__ inline int color_pick(int j, int round, const int2* __restrict__ bodies, const int* __restrict__ btype,
                                 const int* __restrict__ off, const int* __restrict__ adj, const int* state,
                                 int* __restrict__ overflow) {
  const unsigned long long pj = priority(j);
  unsigned long long used0 = 0, used1 = 0;
  bool lose = false;
  const int2 b = bodies[j];
  const bool dyn[2] = {btype[b.x] == 2, btype[b.y] == 2};
  const int side[2] = {b.x, b.y};
  int s0[2], e0[2];
  for (int s = 0; s < 2; ++s) {
    s0[s] = dyn[s] ? off[side[s]] : 0;
    e0[s] = dyn[s] ? off[side[s] + 1] : 0;
  }
  int o[2 * kColorUnroll], so[2 * kColorUnroll];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int t = 0; t < kColorUnroll; ++t) o[s * kColorUnroll + t] = s0[s] + t < e0[s] ? adj[s0[s] + t] : j;
#pragma unroll
  for (int t = 0; t < 2 * kColorUnroll; ++t)
    so[t] = o[t] != j ? __hip_atomic_load(&state[o[t]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0;
#pragma unroll
  for (int t = 0; t < 2 * kColorUnroll; ++t)
    if (o[t] != j) color_look(o[t], so[t], round, pj, lose, used0, used1);
  for (int s = 0; s < 2; ++s)
    for (int e = s0[s] + kColorUnroll; e < e0[s]; ++e) {
      const int oe = adj[e];
      if (oe == j) continue;
      color_look(oe, __hip_atomic_load(&state[oe], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP), round, pj, lose,
                 used0, used1);
    }
  if (lose) return -1;
  if (~used0) return __ffsll(static_cast<long long>(~used0)) - 1;
  if (~used1) return 64 + __ffsll(static_cast<long long>(~used1)) - 1;
  atomicExch(overflow, 1);
  return kMaxColors - 1;
}
// Round 0 of the colouring: a contact keeps its predecessor's colour unless a neighbour of
// higher priority had the same colour (two contacts that persist keep their bodies, so they were
// neighbours before and had different colours; the check only matters after the set changed
// in other ways). Reads only prev_colour, so the outcome does not depend on the launch order.
// Colours at or above `cap` are not kept: those contacts are coloured again first-fit, which
// pulls the colour count back down as the contact graph changes.
// Constraints of sleeping islands are left out of the solve (state -2, never coloured).
// counts[key] += 1 for every thread's key (-1: none), aggregated per block in an LDS table so
// that a key shared by many threads (the root of a large island) takes one global atomic per
// block. Called by every thread of a kBlock-thread block.
__device__ inline void block_count_keys(int key, int* __restrict__ counts) {
  constexpr int H = 2 * kBlock, LG = 9;  // at most kBlock keys: the table stays at most half full
  static_assert((1 << LG) == H, "table size");
  __shared__ int tkey[H];
  __shared__ int tcnt[H];
  for (int e = threadIdx.x; e < H; e += kBlock) {
    tkey[e] = -1;
    tcnt[e] = 0;
  }
  __syncthreads();
  if (key >= 0) {
    int h = static_cast<int>((static_cast<unsigned>(key) * 2654435761u) >> (32 - LG));
    for (;;) {
      const int prev = atomicCAS(&tkey[h], -1, key);
      if (prev == -1 || prev == key) break;
      h = (h + 1) & (H - 1);
    }
    atomicAdd(&tcnt[h], 1);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < H; e += kBlock)
    if (tcnt[e]) atomicAdd(&counts[tkey[e]], tcnt[e]);
}

// Block-aggregated scatter: every thread with a key >= 0 gets a position in its key's range,
// off[key] + (elements of that key placed before it), with one global atomic on cursor[key] per
// distinct key in the block (ranks inside the block from LDS atomics; the order among equal keys
// is free). Called by every thread of a kBlock-thread block; returns -1 for key < 0.
__device__ inline int block_place_keys(int key, const int* __restrict__ off, int* __restrict__ cursor) {
  constexpr int H = 2 * kBlock, LG = 9;
  __shared__ int tkey[H];
  __shared__ int tcnt[H];
  __shared__ int tbase[H];
  for (int e = threadIdx.x; e < H; e += kBlock) {
    tkey[e] = -1;
    tcnt[e] = 0;
  }
  __syncthreads();
  int h = 0, rank = 0;
  if (key >= 0) {
    h = static_cast<int>((static_cast<unsigned>(key) * 2654435761u) >> (32 - LG));
    for (;;) {
      const int prev = atomicCAS(&tkey[h], -1, key);
      if (prev == -1 || prev == key) break;
      h = (h + 1) & (H - 1);
    }
    rank = atomicAdd(&tcnt[h], 1);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < H; e += kBlock)
    if (tcnt[e]) tbase[e] = off[tkey[e]] + atomicAdd(&cursor[tkey[e]], tcnt[e]);
  __syncthreads();
  return key >= 0 ? tbase[h] + rank : -1;
}

__global__ void __launch_bounds__(kBlock) k_color_keep(const int* __restrict__ ncp, const int2* __restrict__ bodies,
                                                       const int* __restrict__ btype, const int* __restrict__ off,
                                                       const int* __restrict__ adj, const int* __restrict__ prev_colour,
                                                       int cap, IslandView iv, int* __restrict__ state,
                                                       int* __restrict__ isl_count,
                                                       int* __restrict__ croot) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  bool kept = false, left_out = false;
  if (j < *ncp && iv.isl_awake) {
    const int2 b = bodies[j];
    left_out = iv.asleep(btype[b.x] != 0 ? b.x : b.y);
    if (left_out) state[j] = -2;
  }
  // each constraint's island (the island solve's bins are made from these and the counts below)
  int root = -1;
  if (j < *ncp && !left_out && isl_count) {
    const int2 b = bodies[j];
    root = iv.parent[btype[b.x] != 0 ? b.x : b.y];
    croot[j] = root;
  }
  if (j < *ncp && !left_out) {
    const int pc = prev_colour[j];
    if (pc >= 0 && pc < cap) {
      kept = true;
      const unsigned long long pj = priority(j);
      const int2 b = bodies[j];
      const int side[2] = {b.x, b.y};
      int s0[2], e0[2];
      for (int s = 0; s < 2; ++s) {
        const bool dyn = btype[side[s]] == 2;
        s0[s] = dyn ? off[side[s]] : 0;
        e0[s] = dyn ? off[side[s] + 1] : 0;
      }
      // as in color_pick: the first kColorUnroll entries per body with independent loads
      int o[2 * kColorUnroll], po[2 * kColorUnroll];
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < kColorUnroll; ++t) o[s * kColorUnroll + t] = s0[s] + t < e0[s] ? adj[s0[s] + t] : j;
#pragma unroll
      for (int t = 0; t < 2 * kColorUnroll; ++t) po[t] = o[t] != j ? prev_colour[o[t]] : -1;
#pragma unroll
      for (int t = 0; t < 2 * kColorUnroll; ++t)
        if (po[t] == pc && priority(o[t]) > pj) kept = false;
      for (int s = 0; s < 2; ++s)
        for (int e = s0[s] + kColorUnroll; e < e0[s]; ++e) {
          const int oe = adj[e];
          if (oe != j && prev_colour[oe] == pc && priority(oe) > pj) kept = false;
        }
    }
    state[j] = kept ? pc : -1;  // round 0
  }
  if (isl_count) block_count_keys(root, isl_count);  // (uniform: every thread of the block calls it)
}

__global__ void __launch_bounds__(kBlock) k_color_round(const int* __restrict__ ncp, int round,
                                                        const int2* __restrict__ bodies, const int* __restrict__ btype,
                                                        const int* __restrict__ off, const int* __restrict__ adj,
                                                        int* state,
                                                        int* __restrict__ overflow) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  int c = -1;
  if (j < *ncp && __hip_atomic_load(&state[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == -1) {
    c = color_pick(j, round, bodies, btype, off, adj, state, overflow);
    if (c >= 0) __hip_atomic_store(&state[j], (round << 8) | c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

// Colour histogram and scatter with block-aggregated atomics: ranks inside a block come from
// LDS atomics and each block reserves one range per colour, so the ~10 global counters see one
// atomic per block instead of one per constraint. The order inside a colour is free (its
// constraints are independent), so no stable ranking is needed.
// count[kMaxColors]: the last round in which a constraint was coloured
// *uncoloured: constraints still without a colour (state -1), which the host reads to decide
// whether more colouring rounds are needed
__global__ void __launch_bounds__(kBlock) k_color_hist(const int* __restrict__ ncp, const int* __restrict__ state,
                                                       int* __restrict__ count, int* __restrict__ uncoloured) {
  __shared__ int h[kMaxColors];
  __shared__ int last_round, unc;
  for (int c = threadIdx.x; c < kMaxColors; c += blockDim.x) h[c] = 0;
  if (threadIdx.x == 0) {
    last_round = 0;
    unc = 0;
  }
  __syncthreads();
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < *ncp) {
    const int st = state[j];
    if (st >= 0) {
      atomicAdd(&h[st & 0xff], 1);
      atomicMax(&last_round, st >> 8);
    } else if (st == -1) {
      atomicAdd(&unc, 1);
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < kMaxColors; c += blockDim.x)
    if (h[c]) atomicAdd(&count[c], h[c]);
  if (threadIdx.x == 0 && last_round > 0) atomicMax(&count[kMaxColors], last_round);
  if (threadIdx.x == 0 && unc > 0) atomicAdd(uncoloured, unc);
}

// map: per colour id its base offset in the colour order ([0, kMaxColors)) and its launch
// index ([kMaxColors, 2 kMaxColors)); the launch index is stored per pair for the next step.
__global__ void __launch_bounds__(kBlock) k_color_scatter(int n, const int* __restrict__ state, const int* __restrict__ map,
                                                          int* __restrict__ cursor, int* __restrict__ order,
                                                          int* __restrict__ inv_order, const int* __restrict__ list,
                                                          int* __restrict__ pair_colour) {
  __shared__ int h[kMaxColors];
  __shared__ int base[kMaxColors];
  for (int c = threadIdx.x; c < kMaxColors; c += blockDim.x) h[c] = 0;
  __syncthreads();
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int c = -1, r = 0;
  if (j < n && state[j] >= 0) {  // left-out constraints (sleeping islands) have no colour
    c = state[j] & 0xff;
    r = atomicAdd(&h[c], 1);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kMaxColors; k += blockDim.x)
    if (h[k]) base[k] = map[k] + atomicAdd(&cursor[k], h[k]);
  __syncthreads();
  if (c >= 0) {
    order[base[c] + r] = j;
    inv_order[j] = base[c] + r;
    pair_colour[list[j]] = map[kMaxColors + c];
  }
}

// ---- constraint setup ---------------------------------------------------------------------------
// B2contactSolver::new + initialize_velocity_constraints for constraint i (colour order):
// src/private/dynamics/b2_contact_solver_private.rs:20-226
struct InitArgs {
  int n;
  const int* order;
  const int* list;
  const int2* ab;
  const int* fbody;
  const int4* fshape;
  const float4* fmat;
  const int* btype;
  const int* parent;
  const float4* bpos;
  const float4* bvel;
  const float4* bmass;
  const int* hdr;
  const float4* geo;
  const float4* pt0;
  const float4* pt1;
  float dt_ratio;
  int warm;
  int block_solve;
  // outputs
  int* k_contact;
  int2* k_bodies;
  int4* k_info;
  float4 *vnf, *vm, *vr0, *vr1, *vms;
  float2* vbias;
  float4 *K, *NM, *lam, *plp, *pgeo, *plc;
  float2* prad;
};

__global__ void __launch_bounds__(kBlock) k_init_constraints(InitArgs a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  int j = a.order[i];
  int k = a.list[j];
  int2 f = a.ab[k];
  int bA = a.fbody[f.x], bB = a.fbody[f.y];
  float4 matA = a.fmat[f.x], matB = a.fmat[f.y];
  // mixing laws, src/b2_contact.rs:29-46
  float friction = sqrtf(matA.x * matB.x);
  float restitution = matA.y > matB.y ? matA.y : matB.y;
  float threshold = matA.z < matB.z ? matA.z : matB.z;
  float tangent_speed = 0.0f;
  float rA = __int_as_float(a.fshape[f.x].z), rB = __int_as_float(a.fshape[f.y].z);
  int h = a.hdr[k];
  int mtype = h & 0xff, mcount = h >> 8;
  float4 g = a.geo[k], p0 = a.pt0[k], p1 = a.pt1[k];
  float4 msA = a.bmass[bA], msB = a.bmass[bB];
  float mA = msA.x, iA = msA.y, mB = msB.x, iB = msB.y;
  V2 lcA = v2(msA.z, msA.w), lcB = v2(msB.z, msB.w);
  float4 pA = a.bpos[bA], pB = a.bpos[bB];
  float4 vA = a.bvel[bA], vB = a.bvel[bB];
  int tA = a.btype[bA], tB = a.btype[bB];
  int island = tA != 0 ? a.parent[bA] : a.parent[bB];
  float4 lam = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (a.warm) {
    lam.x = a.dt_ratio * p0.z;
    lam.y = a.dt_ratio * p0.w;
    if (mcount > 1) {
      lam.z = a.dt_ratio * p1.z;
      lam.w = a.dt_ratio * p1.w;
    }
  }
  Xf xfA, xfB;
  xfA.q = rot(pA.z);
  xfB.q = rot(pB.z);
  xfA.p = v2(pA.x, pA.y) - mul(xfA.q, lcA);
  xfB.p = v2(pB.x, pB.y) - mul(xfB.q, lcB);
  V2 lp[2] = {v2(p0.x, p0.y), v2(p1.x, p1.y)};
  WorldManifold wm;
  world_manifold(&wm, mtype, mcount, v2(g.x, g.y), v2(g.z, g.w), lp, xfA, rA, xfB, rB);
  V2 normal = wm.normal;
  V2 cA = v2(pA.x, pA.y), cB = v2(pB.x, pB.y);
  V2 va = v2(vA.x, vA.y), vb = v2(vB.x, vB.y);
  float4 r[2];
  float nm[2] = {0.0f, 0.0f}, tm[2] = {0.0f, 0.0f}, bias[2] = {0.0f, 0.0f};
  r[0] = r[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int p = 0; p < mcount; ++p) {
    V2 ra = wm.points[p] - cA;
    V2 rb = wm.points[p] - cB;
    r[p] = make_float4(ra.x, ra.y, rb.x, rb.y);
    float rnA = cross(ra, normal), rnB = cross(rb, normal);
    float kNormal = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
    nm[p] = kNormal > 0.0f ? 1.0f / kNormal : 0.0f;
    V2 tangent = cross(normal, 1.0f);
    float rtA = cross(ra, tangent), rtB = cross(rb, tangent);
    float kTangent = mA + mB + iA * rtA * rtA + iB * rtB * rtB;
    tm[p] = kTangent > 0.0f ? 1.0f / kTangent : 0.0f;
    float vRel = dot(normal, ((vb + cross(vB.z, rb)) - va) - cross(vA.z, ra));
    if (vRel < -threshold) bias[p] = -restitution * vRel;
  }
  int vcount = mcount;
  float4 K = make_float4(0.0f, 0.0f, 0.0f, 0.0f), NM = K;
  if (mcount == 2 && a.block_solve) {
    V2 r1a = v2(r[0].x, r[0].y), r1b = v2(r[0].z, r[0].w), r2a = v2(r[1].x, r[1].y), r2b = v2(r[1].z, r[1].w);
    float rn1A = cross(r1a, normal), rn1B = cross(r1b, normal);
    float rn2A = cross(r2a, normal), rn2B = cross(r2b, normal);
    float k11 = mA + mB + iA * rn1A * rn1A + iB * rn1B * rn1B;
    float k22 = mA + mB + iA * rn2A * rn2A + iB * rn2B * rn2B;
    float k12 = mA + mB + iA * rn1A * rn2A + iB * rn1B * rn2B;
    if (k11 * k11 < 1000.0f * (k11 * k22 - k12 * k12)) {
      K = make_float4(k11, k12, k12, k22);
      // B2Mat22::get_inverse (src/b2_math.rs:261-274)
      float det = k11 * k22 - k12 * k12;
      if (det != 0.0f) det = 1.0f / det;
      NM = make_float4(det * k22, -det * k12, -det * k12, det * k11);
    } else {
      vcount = 1;
    }
  }
  a.k_contact[i] = k;
  a.k_bodies[i] = make_int2(bA, bB);
  a.k_info[i] = make_int4(mtype | ((tA == 2) << 8) | ((tB == 2) << 9), mcount, vcount, island);
  a.vnf[i] = make_float4(normal.x, normal.y, friction, tangent_speed);
  a.vm[i] = make_float4(mA, iA, mB, iB);
  a.vr0[i] = r[0];
  a.vr1[i] = r[1];
  a.vms[i] = make_float4(nm[0], tm[0], nm[1], tm[1]);
  a.vbias[i] = make_float2(bias[0], bias[1]);
  a.K[i] = K;
  a.NM[i] = NM;
  a.lam[i] = lam;
  a.plp[i] = make_float4(p0.x, p0.y, p1.x, p1.y);
  a.pgeo[i] = g;
  a.plc[i] = make_float4(lcA.x, lcA.y, lcB.x, lcB.y);
  a.prad[i] = make_float2(rA, rB);
}

// ---- velocity passes --------------------------------------------------------------------------
struct VelArgs {
  const int2* bodies;
  const int4* info;
  const float4 *vnf, *vm, *vr0, *vr1, *vms;
  const float2* vbias;
  const float4 *K, *NM;
  float4* lam;
  float4* vel;
};

// B2contactSolver::warm_start (src/private/dynamics/b2_contact_solver_private.rs:228-266) as
// one pass over dynamic bodies: a body applies the warm-start impulses of its constraints in
// colour order (ascending colour-sorted index), the order in which a sweep colour by colour
// applies them to it, with the same operations, so the velocities are the same bit for bit (two
// constraints of one colour never share a dynamic body). One launch instead of one per colour.
// one constraint's warm-start data as seen from body b (its side, point count, masses, normal,
// impulses and anchors)
struct WarmRec {
  float4 m, nf, lam, r0, r1;
  int side_a, count;
};
__device__ inline WarmRec warm_load(int b, int i, const VelArgs& a) {
  WarmRec r;
  r.side_a = a.bodies[i].x == b;
  r.count = a.info[i].z;
  r.m = a.vm[i];
  r.nf = a.vnf[i];
  r.lam = a.lam[i];
  r.r0 = a.vr0[i];
  r.r1 = a.vr1[i];
  return r;
}
__device__ inline void warm_apply(const WarmRec& c, V2& v, float& w) {
  const V2 normal = v2(c.nf.x, c.nf.y);
  const V2 tangent = cross(normal, 1.0f);
  for (int p = 0; p < c.count; ++p) {
    const float4 r = p == 0 ? c.r0 : c.r1;
    const float ni = p == 0 ? c.lam.x : c.lam.z, ti = p == 0 ? c.lam.y : c.lam.w;
    const V2 P = ni * normal + ti * tangent;
    if (c.side_a) {
      w -= c.m.y * cross(v2(r.x, r.y), P);
      v = v - c.m.x * P;
    } else {
      w += c.m.w * cross(v2(r.z, r.w), P);
      v = v + c.m.z * P;
    }
  }
}
// Up to kWarmSort constraints per body are sorted in registers (odd-even transposition network)
// and their records loaded kWarmBatch at a time with independent loads before they are applied
// in order; a body with more constraints takes the next larger index each time.
static constexpr int kWarmSort = 16, kWarmBatch = 8;
__global__ void __launch_bounds__(kBlock) k_warm_start_bodies(int nb, const int* __restrict__ btype,
                                                              const int* __restrict__ off, const int* __restrict__ adj,
                                                              const int* __restrict__ inv_order, VelArgs a,
                                                              IslandView iv) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb || btype[b] != 2 || iv.asleep(b)) return;
  const int s0 = off[b], e0 = off[b + 1], n = e0 - s0;
  if (n == 0) return;
  const float4 vb = a.vel[b];
  V2 v = v2(vb.x, vb.y);
  float w = vb.z;
  if (n <= kWarmSort) {
    int k[kWarmSort];
#pragma unroll
    for (int t = 0; t < kWarmSort; ++t) k[t] = t < n ? inv_order[adj[s0 + t]] : INT_MAX;
#pragma unroll
    for (int r = 0; r < kWarmSort; ++r)
#pragma unroll
      for (int t = r & 1; t + 1 < kWarmSort; t += 2) {
        const int lo = min(k[t], k[t + 1]), hi = max(k[t], k[t + 1]);
        k[t] = lo;
        k[t + 1] = hi;
      }
#pragma unroll
    for (int h = 0; h < kWarmSort; h += kWarmBatch) {
      if (h >= n) break;
      WarmRec rec[kWarmBatch];
#pragma unroll
      for (int t = 0; t < kWarmBatch; ++t) rec[t] = warm_load(b, k[h + t] != INT_MAX ? k[h + t] : k[0], a);
#pragma unroll
      for (int t = 0; t < kWarmBatch; ++t)
        if (k[h + t] != INT_MAX) warm_apply(rec[t], v, w);
    }
  } else {
    int last = -1;
    for (int c = 0; c < n; ++c) {
      int next = INT_MAX;
      for (int e = s0; e < e0; ++e) {
        const int i = inv_order[adj[e]];
        if (i > last && i < next) next = i;
      }
      warm_apply(warm_load(b, next, a), v, w);
      last = next;
    }
  }
  a.vel[b] = make_float4(v.x, v.y, w, 0.0f);
}

// B2contactSolver::solve_velocity_constraints for one constraint
// (src/private/dynamics/b2_contact_solver_private.rs:268-583)
__device__ inline void apply_normal_pair(V2& va, float& wa, V2& vb, float& wb, float mA, float iA, float mB, float iB,
                                         V2 normal, V2 r1a, V2 r2a, V2 r1b, V2 r2b, V2 d) {
  V2 P1 = d.x * normal, P2 = d.y * normal;
  va = va - mA * (P1 + P2);
  wa -= iA * (cross(r1a, P1) + cross(r2a, P2));
  vb = vb + mB * (P1 + P2);
  wb += iB * (cross(r1b, P1) + cross(r2b, P2));
}

// The constraint's records and the velocities (v.x, v.y, w, 0) of its two bodies in, its new
// impulses and velocities out. kload(K, NM) supplies the block solver's K and normal mass; it is
// called on that path only.
template <class KLoad>
__device__ inline void velocity_core(const float4 m, const float4 nf, float4& lam, const float4 ms, const float4 r0,
                                     const float4 r1, const int vcount, const float2 bias, const int block_solve,
                                     KLoad kload, float4& vA, float4& vB) {
  const float mA = m.x, iA = m.y, mB = m.z, iB = m.w;
  V2 va = v2(vA.x, vA.y), vb = v2(vB.x, vB.y);
  float wa = vA.z, wb = vB.z;
  V2 normal = v2(nf.x, nf.y);
  V2 tangent = cross(normal, 1.0f);
  const float friction = nf.z;
  V2 r1a = v2(r0.x, r0.y), r1b = v2(r0.z, r0.w);
  V2 r2a = v2(r1.x, r1.y), r2b = v2(r1.z, r1.w);
  // tangent constraints
  for (int p = 0; p < vcount; ++p) {
    V2 ra = p == 0 ? r1a : r2a, rb = p == 0 ? r1b : r2b;
    float tmass = p == 0 ? ms.y : ms.w;
    float nimp = p == 0 ? lam.x : lam.z;
    float timp = p == 0 ? lam.y : lam.w;
    V2 dv = ((vb + cross(wb, rb)) - va) - cross(wa, ra);
    float vt = dot(dv, tangent) - nf.w;
    float lambda = tmass * (-vt);
    float maxFriction = friction * nimp;
    float newImpulse = bclamp(timp + lambda, -maxFriction, maxFriction);
    lambda = newImpulse - timp;
    if (p == 0) lam.y = newImpulse; else lam.w = newImpulse;
    V2 P = lambda * tangent;
    va = va - mA * P;
    wa -= iA * cross(ra, P);
    vb = vb + mB * P;
    wb += iB * cross(rb, P);
  }
  if (vcount == 1 || !block_solve) {
    for (int p = 0; p < vcount; ++p) {
      V2 ra = p == 0 ? r1a : r2a, rb = p == 0 ? r1b : r2b;
      float nmass = p == 0 ? ms.x : ms.z;
      float vbias = p == 0 ? bias.x : bias.y;
      float nimp = p == 0 ? lam.x : lam.z;
      V2 dv = ((vb + cross(wb, rb)) - va) - cross(wa, ra);
      float vn = dot(dv, normal);
      float lambda = -nmass * (vn - vbias);
      float newImpulse = bmax(nimp + lambda, 0.0f);
      lambda = newImpulse - nimp;
      if (p == 0) lam.x = newImpulse; else lam.z = newImpulse;
      V2 P = lambda * normal;
      va = va - mA * P;
      wa -= iA * cross(ra, P);
      vb = vb + mB * P;
      wb += iB * cross(rb, P);
    }
  } else {
    float4 K, NM;  // (ex.x, ex.y, ey.x, ey.y)
    kload(K, NM);
    V2 aa = v2(lam.x, lam.z);
    V2 dv1 = ((vb + cross(wb, r1b)) - va) - cross(wa, r1a);
    V2 dv2 = ((vb + cross(wb, r2b)) - va) - cross(wa, r2a);
    float vn1 = dot(dv1, normal);
    float vn2 = dot(dv2, normal);
    V2 b = v2(vn1 - bias.x, vn2 - bias.y);
    V2 Ka = v2(K.x * aa.x + K.z * aa.y, K.y * aa.x + K.w * aa.y);
    b = b - Ka;
    for (;;) {
      // case 1: vn = 0
      V2 x = -v2(NM.x * b.x + NM.z * b.y, NM.y * b.x + NM.w * b.y);
      if (x.x >= 0.0f && x.y >= 0.0f) {
        apply_normal_pair(va, wa, vb, wb, mA, iA, mB, iB, normal, r1a, r2a, r1b, r2b, x - aa);
        lam.x = x.x;
        lam.z = x.y;
        break;
      }
      // case 2: vn1 = 0 and x2 = 0
      x.x = -ms.x * b.x;
      x.y = 0.0f;
      vn2 = K.y * x.x + b.y;
      if (x.x >= 0.0f && vn2 >= 0.0f) {
        apply_normal_pair(va, wa, vb, wb, mA, iA, mB, iB, normal, r1a, r2a, r1b, r2b, x - aa);
        lam.x = x.x;
        lam.z = x.y;
        break;
      }
      // case 3: vn2 = 0 and x1 = 0
      x.x = 0.0f;
      x.y = -ms.z * b.y;
      vn1 = K.z * x.y + b.x;
      if (x.y >= 0.0f && vn1 >= 0.0f) {
        apply_normal_pair(va, wa, vb, wb, mA, iA, mB, iB, normal, r1a, r2a, r1b, r2b, x - aa);
        lam.x = x.x;
        lam.z = x.y;
        break;
      }
      // case 4: x1 = 0 and x2 = 0
      x.x = 0.0f;
      x.y = 0.0f;
      vn1 = b.x;
      vn2 = b.y;
      if (vn1 >= 0.0f && vn2 >= 0.0f) {
        apply_normal_pair(va, wa, vb, wb, mA, iA, mB, iB, normal, r1a, r2a, r1b, r2b, x - aa);
        lam.x = x.x;
        lam.z = x.y;
        break;
      }
      break;  // no solution: give up
    }
  }
  vA = make_float4(va.x, va.y, wa, 0.0f);
  vB = make_float4(vb.x, vb.y, wb, 0.0f);
}

__device__ inline void solve_velocity_one(int i, int block_solve, const VelArgs& a) {
  const int2 bd = a.bodies[i];
  const int4 info = a.info[i];
  const float4 m = a.vm[i];
  const float4 nf = a.vnf[i];
  float4 lam = a.lam[i];
  const float4 ms = a.vms[i];
  const float4 r0 = a.vr0[i];
  float4 vA = a.vel[bd.x], vB = a.vel[bd.y];
  const int vcount = info.z;
  float4 r1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (vcount == 2) r1 = a.vr1[i];
  const float2 bias = a.vbias[i];
  velocity_core(m, nf, lam, ms, r0, r1, vcount, bias, block_solve,
                [&](float4& K, float4& NM) {
                  K = a.K[i];
                  NM = a.NM[i];
                },
                vA, vB);
  a.lam[i] = lam;
  if (info.x & 0x100) a.vel[bd.x] = vA;
  if (info.x & 0x200) a.vel[bd.y] = vB;
}
__global__ void __launch_bounds__(kBlock) k_solve_velocity(int begin, int end, int block_solve, VelArgs a) {
  int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < end) solve_velocity_one(i, block_solve, a);
}

// The small last colours of first-fit colouring (each at most kTailMax constraints) are solved by
// one workgroup in one launch: colour after colour, with a workgroup barrier between them, which
// is the same sequence as one launch per colour (the barrier makes a colour's velocity writes
// visible to the workgroup's threads before the next colour reads them).
static constexpr int kTailBlock = 1024, kTailMax = 512;
// off holds ng ranges per colour; the tail solves the last one (ng 1: whole colours; with island
// bins, the constraints of the islands too large for a bin)
__global__ void __launch_bounds__(kTailBlock) k_solve_velocity_tail(int c0, int c1, const int* __restrict__ off,
                                                                    int ng, int block_solve, VelArgs a) {
  for (int c = c0; c < c1; ++c) {
    const int b0 = off[c * ng + ng - 1], b1 = off[(c + 1) * ng];
    for (int i = b0 + threadIdx.x; i < b1; i += kTailBlock) solve_velocity_one(i, block_solve, a);
    __syncthreads();
  }
}

// store_impulses (src/private/dynamics/b2_contact_solver_private.rs:585-598)
__device__ inline void store_one(int i, const int* __restrict__ kc, const int4* __restrict__ info,
                                 const float4* __restrict__ lam, float4* __restrict__ pt0, float4* __restrict__ pt1) {
  int k = kc[i];
  int vcount = info[i].z;
  float4 l = lam[i];
  float4 p = pt0[k];
  p.z = l.x;
  p.w = l.y;
  pt0[k] = p;
  if (vcount > 1) {
    float4 q = pt1[k];
    q.z = l.z;
    q.w = l.w;
    pt1[k] = q;
  }
}
__global__ void __launch_bounds__(kBlock) k_store_impulses(int n, const int* __restrict__ kc, const int4* __restrict__ info,
                                                           const float4* __restrict__ lam, float4* __restrict__ pt0,
                                                           float4* __restrict__ pt1) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) store_one(i, kc, info, lam, pt0, pt1);
}
// the impulse store (per constraint) and the position integration (per body) in one launch:
// neither reads what the other writes
__global__ void __launch_bounds__(kBlock) k_store_and_integrate(int nc, const int* __restrict__ kc,
                                                                const int4* __restrict__ info,
                                                                const float4* __restrict__ lam, float4* __restrict__ pt0,
                                                                float4* __restrict__ pt1, int nb,
                                                                const int* __restrict__ btype, float4* __restrict__ pos,
                                                                float4* __restrict__ vel, float h, IslandView iv) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nc) store_one(t, kc, info, lam, pt0, pt1);
  if (t < nb) integrate_position_one(t, btype, pos, vel, h, iv);
}

// ---- position passes ----------------------------------------------------------------------------
struct PosArgs {
  const int2* bodies;
  const int4* info;
  const float4 *vm, *plp, *pgeo, *plc;
  const float2* prad;
  float4* pos;
  int* violated;  // per island root: a contact of the island is below -3 linear slop this iteration
  int* done;      // per island root: iterations run once it stopped (read by the per-colour kernels)
};

// The constraint's position records and the positions (c.x, c.y, angle, 0) of its two bodies in,
// the corrected positions out; returns the smallest separation seen (0 for count 0).
__device__ inline float position_core(const int type, const int count, const float4 m, const float4 lp,
                                      const float4 g, const float4 lc, const float2 rad, float4& PA, float4& PB) {
  const float mA = m.x, iA = m.y, mB = m.z, iB = m.w;
  V2 lcA = v2(lc.x, lc.y), lcB = v2(lc.z, lc.w);
  V2 cA = v2(PA.x, PA.y), cB = v2(PB.x, PB.y);
  float aA = PA.z, aB = PB.z;
  V2 ln = v2(g.x, g.y), lpt = v2(g.z, g.w);
  float minSep = 0.0f;
  for (int p = 0; p < count; ++p) {
    Xf xfA, xfB;
    xfA.q = rot(aA);
    xfB.q = rot(aB);
    xfA.p = cA - mul(xfA.q, lcA);
    xfB.p = cB - mul(xfB.q, lcB);
    V2 normal, point;
    float separation;
    V2 lpp = p == 0 ? v2(lp.x, lp.y) : v2(lp.z, lp.w);
    if (type == kCircles) {
      V2 pA = mul(xfA, lpt);
      V2 pB = mul(xfB, v2(lp.x, lp.y));
      normal = pB - pA;
      normalize(&normal);
      point = 0.5f * (pA + pB);
      separation = dot(pB - pA, normal) - rad.x - rad.y;
    } else if (type == kFaceA) {
      normal = mul(xfA.q, ln);
      V2 plane = mul(xfA, lpt);
      V2 clip = mul(xfB, lpp);
      separation = dot(clip - plane, normal) - rad.x - rad.y;
      point = clip;
    } else {
      normal = mul(xfB.q, ln);
      V2 plane = mul(xfB, lpt);
      V2 clip = mul(xfA, lpp);
      separation = dot(clip - plane, normal) - rad.x - rad.y;
      point = clip;
      normal = -normal;
    }
    V2 rA = point - cA, rB = point - cB;
    minSep = bmin(minSep, separation);
    float C = bclamp(kBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, 0.0f);
    float rnA = cross(rA, normal), rnB = cross(rB, normal);
    float K = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
    float impulse = K > 0.0f ? -C / K : 0.0f;
    V2 P = impulse * normal;
    cA = cA - mA * P;
    aA -= iA * cross(rA, P);
    cB = cB + mB * P;
    aB += iB * cross(rB, P);
  }
  PA = make_float4(cA.x, cA.y, aA, 0.0f);
  PB = make_float4(cB.x, cB.y, aB, 0.0f);
  return minSep;
}

// B2contactSolver::solve_position_constraints for one constraint
// (src/private/dynamics/b2_contact_solver_private.rs:600-730)
// Lanes past the colour's end pass live = false: they read the colour's first constraint and
// write nothing. All loads are issued up front: the constraint's records, then the body positions
// and the island flag (written by earlier launches only) side by side.
__device__ inline void solve_position_one(int i, bool live, int begin, const PosArgs& a) {
  const int ii = live ? i : begin;
  int4 info = a.info[ii];
  const int2 bd = a.bodies[ii];
  const float4 m = a.vm[ii];
  const float4 lp = a.plp[ii], g = a.pgeo[ii], lc = a.plc[ii];
  const float2 rad = a.prad[ii];
  float4 PA = a.pos[bd.x], PB = a.pos[bd.y];
  const bool active = live && a.done[info.w] == 0;
  if (!active) info.y = 0;
  const float minSep = position_core(info.x & 0xff, info.y, m, lp, g, lc, rad, PA, PB);
  if (active && (info.x & 0x100)) a.pos[bd.x] = PA;
  if (active && (info.x & 0x200)) a.pos[bd.y] = PB;
  // The island may stop iterating only if no contact's separation is below -3 linear slop
  // (minSeparation >= -3 * slop over the island's contacts, b2_contact_solver_private.rs); a
  // violating constraint marks its island with a plain store (every writer stores the same value).
  if (active && !(minSep >= -3.0f * kLinearSlop)) a.violated[info.w] = 1;
}
__global__ void __launch_bounds__(kBlock) k_solve_position(int begin, int end, PosArgs a) {
  int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  solve_position_one(i, i < end, begin, a);
}

// The overflow colour: a dynamic body with more contacts than the colour masks hold (kMaxColors)
// gives its last contacts the highest colour, where two constraints may share a body. That colour
// is solved one constraint after another on one thread: Gauss-Seidel in its order, as the
// reference solves every constraint (b2_contact_solver_private.rs:268-730). Only worlds with such
// a body take this path.
__global__ void __launch_bounds__(64) k_solve_velocity_serial(int begin, int end, int block_solve, VelArgs a) {
  if (threadIdx.x != 0) return;
  for (int i = begin; i < end; ++i) solve_velocity_one(i, block_solve, a);
}
__global__ void __launch_bounds__(64) k_solve_position_serial(int begin, int end, PosArgs a) {
  if (threadIdx.x != 0) return;
  for (int i = begin; i < end; ++i) solve_position_one(i, true, begin, a);
}
__global__ void __launch_bounds__(kTailBlock) k_solve_position_tail(int c0, int c1, const int* __restrict__ off,
                                                                    int ng, PosArgs a) {
  for (int c = c0; c < c1; ++c) {
    const int b0 = off[c * ng + ng - 1], b1 = off[(c + 1) * ng];
    for (int i = b0 + threadIdx.x; i < b1; i += kTailBlock) solve_position_one(i, true, b0, a);
    __syncthreads();
  }
}

__device__ inline void iteration_end_root(int r, int iteration, int* __restrict__ violated, int* __restrict__ done) {
  if (!violated[r]) done[r] = iteration + 1;
  violated[r] = 0;
}
// end of a position iteration: islands whose error is small stop iterating
// (src/private/dynamics/b2_island_private.rs:257-273). A finished island's flag holds the number
// of iterations it ran (the adjoint's tape reads it).
__device__ inline void iteration_end_one(int b, int iteration, const int* __restrict__ btype,
                                         const int* __restrict__ parent, int* __restrict__ violated,
                                         int* __restrict__ done, const int* __restrict__ isl_awake) {
  if (btype[b] == 0 || parent[b] != b || done[b] || (isl_awake && !isl_awake[b])) return;
  iteration_end_root(b, iteration, violated, done);
}
// the last position iteration's end and the copy back in one launch (both per body; finalize
// reads only positions, the iteration end only island flags)
struct IterationEndArgs {
  int iteration;  // -1: no iteration ends here
  const int* parent;
  int* violated;
  int* done;
  const int* isl_awake;
};
__global__ void __launch_bounds__(kBlock) k_finalize_end(int nb, const int* __restrict__ btype,
                                                         const float4* __restrict__ pos, const float4* __restrict__ bmass,
                                                         float4* __restrict__ xf, IslandView iv, IterationEndArgs e) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  if (e.iteration >= 0) iteration_end_one(b, e.iteration, btype, e.parent, e.violated, e.done, e.isl_awake);
  finalize_one(b, btype, pos, bmass, xf, iv);
}
// binned (optional): islands of at most bin_cap constraints, whose iteration ends the island
// solve's position kernel has run, are left out
__global__ void __launch_bounds__(kBlock) k_position_iteration_end(int nb, int iteration, const int* __restrict__ btype,
                                                                   const int* __restrict__ parent,
                                                                   int* __restrict__ violated, int* __restrict__ done,
                                                                   const int* __restrict__ isl_awake,
                                                                   const int* __restrict__ binned, int bin_cap) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  if (binned && binned[b] > 0 && binned[b] <= bin_cap) return;
  iteration_end_one(b, iteration, btype, parent, violated, done, isl_awake);
}

// ---- per-world solve (batched worlds) ---------------------------------------------------------
// Within each colour the constraints are grouped by world: cw_off[L * nw + w] is the first
// constraint of world w in launch colour L (an exclusive scan over (colour, world), colour-major,
// so cw_off[L * nw] is colour L's first constraint and the order stays colour-major). One
// workgroup then runs all iterations of its world, colour after colour with a workgroup barrier
// between colours -- the sequence of the per-colour launches restricted to one world, whose
// constraints touch no other world's bodies, so the results are the same bit for bit.
static constexpr int kWorldSolveMin = 16;

__device__ inline int constraint_world(int j, const int* __restrict__ list, const int2* __restrict__ ab,
                                       const int* __restrict__ fworld) {
  return fworld[ab[list[j]].x];
}

__global__ void __launch_bounds__(kBlock) k_color_world_hist(int n, const int* __restrict__ state,
                                                             const int* __restrict__ launch, const int* __restrict__ list,
                                                             const int2* __restrict__ ab, const int* __restrict__ fworld,
                                                             int nw, int* __restrict__ count) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = j < n && state[j] >= 0 ? launch[state[j] & 0xff] * nw + constraint_world(j, list, ab, fworld) : -1;
  block_count_keys(g, count);
}

__global__ void __launch_bounds__(kBlock) k_color_world_scatter(int n, const int* __restrict__ state,
                                                                const int* __restrict__ launch,
                                                                const int* __restrict__ list, const int2* __restrict__ ab,
                                                                const int* __restrict__ fworld, int nw,
                                                                const int* __restrict__ off, int* __restrict__ cursor,
                                                                int* __restrict__ order, int* __restrict__ inv_order,
                                                                int* __restrict__ pair_colour) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int L = j < n && state[j] >= 0 ? launch[state[j] & 0xff] : -1;
  const int g = L >= 0 ? L * nw + constraint_world(j, list, ab, fworld) : -1;
  const int pos = block_place_keys(g, off, cursor);
  if (g >= 0) {
    order[pos] = j;
    inv_order[j] = pos;
    pair_colour[list[j]] = L;
  }
}

__global__ void __launch_bounds__(kBlock) k_clear_ints(int n, int* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// the world of this workgroup and its constraint range per colour (in LDS)
template <int B>
__device__ inline int world_ranges(const int* __restrict__ off, int nw, int ncolors, const int* __restrict__ wlist,
                                   int wbase, int2* rng) {
  const int w = wlist[wbase + blockIdx.x];
  for (int c = threadIdx.x; c < ncolors; c += B) rng[c] = make_int2(off[c * nw + w], off[c * nw + w + 1]);
  __syncthreads();
  return w;
}

template <int B>
__global__ void __launch_bounds__(B) k_solve_worlds_velocity(const int* __restrict__ off, int nw, int ncolors,
                                                             int iters, int block_solve, VelArgs a,
                                                             const int* __restrict__ wlist, int wbase) {
  __shared__ int2 rng[kMaxColors];
  world_ranges<B>(off, nw, ncolors, wlist, wbase, rng);
  for (int it = 0; it < iters; ++it)
    for (int c = 0; c < ncolors; ++c) {
      const int2 r = rng[c];
      if (r.x == r.y) continue;  // uniform across the workgroup
      for (int i = r.x + threadIdx.x; i < r.y; i += B) solve_velocity_one(i, block_solve, a);
      __syncthreads();
    }
}

// position iterations with the islands' early exit: after every iteration but the last (whose
// end runs with the copy back), the world's island roots are updated as k_position_iteration_end
// updates all of them
template <int B>
__global__ void __launch_bounds__(B) k_solve_worlds_position(const int* __restrict__ off, int nw, int ncolors,
                                                             int iters, const int* __restrict__ body_off,
                                                             const int* __restrict__ btype, PosArgs a,
                                                             IterationEndArgs e, const int* __restrict__ wlist,
                                                             int wbase) {
  __shared__ int2 rng[kMaxColors];
  const int w = world_ranges<B>(off, nw, ncolors, wlist, wbase, rng);
  const int b0 = body_off[w], b1 = body_off[w + 1];
  for (int it = 0; it < iters; ++it) {
    for (int c = 0; c < ncolors; ++c) {
      const int2 r = rng[c];
      if (r.x == r.y) continue;
      for (int i = r.x + threadIdx.x; i < r.y; i += B) solve_position_one(i, true, r.x, a);
      __syncthreads();
    }
    if (it + 1 < iters) {
      for (int b = b0 + threadIdx.x; b < b1; b += B)
        iteration_end_one(b, it, btype, e.parent, e.violated, e.done, e.isl_awake);
      __syncthreads();
    }
  }
}

// one launch per body-count bucket of worlds, with that bucket's threads per world; a timed launch
// counts its share of the items by its number of worlds
static void launch_worlds_velocity(World* w, long long items, const int* off, int ncolors, int iters, int block_solve,
                                   const VelArgs& a) {
  const int nw = w->n_worlds;
  const int* wl = w->s.world_order.p;
  for (int k = 0; k < 3; ++k) {
    const int n = w->world_bucket_off[k + 1] - w->world_bucket_off[k], wb = w->world_bucket_off[k];
    if (n == 0) continue;
    const long long share = items * n / std::max(nw, 1);
    const int K = B2G_KERNEL_SOLVE_VELOCITY;
    switch (w->world_threads[k]) {
      case 64:
        launch_timed(w, K, share, k_solve_worlds_velocity<64>, dim3(n), dim3(64), off, nw, ncolors, iters, block_solve, a, wl, wb);
        break;
      case 128:
        launch_timed(w, K, share, k_solve_worlds_velocity<128>, dim3(n), dim3(128), off, nw, ncolors, iters, block_solve, a, wl, wb);
        break;
      case 512:
        launch_timed(w, K, share, k_solve_worlds_velocity<512>, dim3(n), dim3(512), off, nw, ncolors, iters, block_solve, a, wl, wb);
        break;
      default:
        launch_timed(w, K, share, k_solve_worlds_velocity<256>, dim3(n), dim3(256), off, nw, ncolors, iters, block_solve, a, wl, wb);
        break;
    }
  }
}

static void launch_worlds_position(World* w, long long items, const int* off, int ncolors, int iters,
                                   const PosArgs& a, const IterationEndArgs& e) {
  const int nw = w->n_worlds;
  const int* wl = w->s.world_order.p;
  const int* bo = w->s.world_body_off.p;
  const int* bt = w->b.type.p;
  for (int k = 0; k < 3; ++k) {
    const int n = w->world_bucket_off[k + 1] - w->world_bucket_off[k], wb = w->world_bucket_off[k];
    if (n == 0) continue;
    const long long share = items * n / std::max(nw, 1);
    const int K = B2G_KERNEL_SOLVE_POSITION;
    switch (w->world_threads[k]) {
      case 64:
        launch_timed(w, K, share, k_solve_worlds_position<64>, dim3(n), dim3(64), off, nw, ncolors, iters, bo, bt, a, e, wl, wb);
        break;
      case 128:
        launch_timed(w, K, share, k_solve_worlds_position<128>, dim3(n), dim3(128), off, nw, ncolors, iters, bo, bt, a, e, wl, wb);
        break;
      case 512:
        launch_timed(w, K, share, k_solve_worlds_position<512>, dim3(n), dim3(512), off, nw, ncolors, iters, bo, bt, a, e, wl, wb);
        break;
      default:
        launch_timed(w, K, share, k_solve_worlds_position<256>, dim3(n), dim3(256), off, nw, ncolors, iters, bo, bt, a, e, wl, wb);
        break;
    }
  }
}

// ---- island solve (island bins) ---------------------------------------------------------------
// Islands share no dynamic body, so any interleaving of their constraints that keeps each
// island's colour order is the same Gauss-Seidel sequence as the per-colour launches. Whole
// islands of at most W = T/2 constraints are packed into bins: island r (its root) goes to bin
// start(r) / W, where start is the exclusive scan of the small islands' counts in root order, so a
// bin holds fewer than 2W = T constraints. One workgroup of T threads solves a bin through all
// velocity (or position) iterations, colour after colour with a workgroup barrier between them;
// each thread holds one constraint in registers and the bin's bodies sit in an LDS hash table
// (body id -> slot), loaded once and written back once. Islands of more than W constraints keep
// the per-colour launches (their constraints are the last group of every colour).
static constexpr int kBinHashPerThread = 4;  // body slots per thread (a bin has < 2T bodies)

// slot of body (or island) id `key` in a linear-probing table of H = 2^lg entries (-1 = empty);
// the table never fills (at most H/2 keys), so the probe ends
__device__ inline int bin_slot(int* keys, int lg, int key) {
  const int mask = (1 << lg) - 1;
  int h = static_cast<int>((static_cast<unsigned>(key) * 2654435761u) >> (32 - lg));
  for (;;) {
    const int prev = atomicCAS(&keys[h], -1, key);
    if (prev == -1 || prev == key) return h;
    h = (h + 1) & mask;
  }
}

// this thread's constraint: the t-th of its bin in colour order (-1: none); rng holds the bin's
// range per colour
__device__ inline int bin_constraint(const int2* rng, int ncolors, int t, int* colour) {
  int base = 0;
  for (int c = 0; c < ncolors; ++c) {
    const int2 r = rng[c];
    const int n = r.y - r.x;
    if (t >= base && t < base + n) {
      *colour = c;
      return r.x + t - base;
    }
    base += n;
  }
  *colour = -1;
  return -1;
}

template <int T>
__device__ inline void bin_ranges(const int* __restrict__ off, int ng, int ncolors, int2* rng) {
  const int g = blockIdx.x;
  for (int c = threadIdx.x; c < ncolors; c += T) rng[c] = make_int2(off[c * ng + g], off[c * ng + g + 1]);
}

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

template <int T>
__global__ void __launch_bounds__(T) k_solve_bins_velocity(const int* __restrict__ off, int ng, int ncolors, int iters,
                                                           int block_solve, VelArgs a) {
  constexpr int H = kBinHashPerThread * T, LG = ilog2(H);
  __shared__ int2 rng[kMaxColors];
  __shared__ int keys[H];
  __shared__ unsigned char dirty[H];
  __shared__ float4 bv[H];
  bin_ranges<T>(off, ng, ncolors, rng);
  for (int e = threadIdx.x; e < H; e += T) {
    keys[e] = -1;
    dirty[e] = 0;
  }
  __syncthreads();
  int mc;
  const int i = bin_constraint(rng, ncolors, threadIdx.x, &mc);
  int2 bd = make_int2(0, 0);
  int4 info = make_int4(0, 0, 0, 0);
  float4 m = {}, nf = {}, lam = {}, ms = {}, r0 = {}, r1 = {}, K = {}, NM = {};
  float2 bias = {};
  int sa = 0, sb = 0;
  if (i >= 0) {
    bd = a.bodies[i];
    info = a.info[i];
    m = a.vm[i];
    nf = a.vnf[i];
    lam = a.lam[i];
    ms = a.vms[i];
    r0 = a.vr0[i];
    if (info.z == 2) r1 = a.vr1[i];
    bias = a.vbias[i];
    if (block_solve && info.z == 2) {
      K = a.K[i];
      NM = a.NM[i];
    }
    sa = bin_slot(keys, LG, bd.x);
    sb = bin_slot(keys, LG, bd.y);
    if (info.x & 0x100) dirty[sa] = 1;
    if (info.x & 0x200) dirty[sb] = 1;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < H; e += T) {
    const int k = keys[e];
    if (k >= 0) bv[e] = a.vel[k];
  }
  __syncthreads();
  for (int it = 0; it < iters; ++it)
    for (int c = 0; c < ncolors; ++c) {
      if (rng[c].x == rng[c].y) continue;  // uniform across the workgroup
      if (mc == c) {
        float4 vA = bv[sa], vB = bv[sb];
        velocity_core(m, nf, lam, ms, r0, r1, info.z, bias, block_solve,
                      [&](float4& k1, float4& k2) {
                        k1 = K;
                        k2 = NM;
                      },
                      vA, vB);
        if (info.x & 0x100) bv[sa] = vA;
        if (info.x & 0x200) bv[sb] = vB;
      }
      __syncthreads();
    }
  if (i >= 0) a.lam[i] = lam;
  for (int e = threadIdx.x; e < H; e += T)
    if (dirty[e]) a.vel[keys[e]] = bv[e];
}

// Position iterations with each island's early exit: the islands' flags (iterations run,
// violated) sit in a second table keyed by island root; after every iteration but the last
// (whose end runs with the copy back) each island of the bin ends its iteration as
// k_position_iteration_end ends it. The flags are written back for that last end.
template <int T>
__global__ void __launch_bounds__(T) k_solve_bins_position(const int* __restrict__ off, int ng, int ncolors, int iters,
                                                           PosArgs a) {
  constexpr int H = kBinHashPerThread * T, LG = ilog2(H);
  constexpr int HR = 2 * T, LGR = ilog2(HR);  // island table: fewer than T roots
  __shared__ int2 rng[kMaxColors];
  __shared__ int keys[H];
  __shared__ unsigned char dirty[H];
  __shared__ float4 bp[H];
  __shared__ int rkeys[HR];
  __shared__ int rdone[HR];
  __shared__ unsigned char rviol[HR];
  bin_ranges<T>(off, ng, ncolors, rng);
  for (int e = threadIdx.x; e < H; e += T) {
    keys[e] = -1;
    dirty[e] = 0;
  }
  for (int e = threadIdx.x; e < HR; e += T) rkeys[e] = -1;
  __syncthreads();
  int mc;
  const int i = bin_constraint(rng, ncolors, threadIdx.x, &mc);
  int2 bd = make_int2(0, 0);
  int4 info = make_int4(0, 0, 0, 0);
  float4 m = {}, lp = {}, g = {}, lc = {};
  float2 rad = {};
  int sa = 0, sb = 0, sr = 0;
  if (i >= 0) {
    bd = a.bodies[i];
    info = a.info[i];
    m = a.vm[i];
    lp = a.plp[i];
    g = a.pgeo[i];
    lc = a.plc[i];
    rad = a.prad[i];
    sa = bin_slot(keys, LG, bd.x);
    sb = bin_slot(keys, LG, bd.y);
    sr = bin_slot(rkeys, LGR, info.w);
    if (info.x & 0x100) dirty[sa] = 1;
    if (info.x & 0x200) dirty[sb] = 1;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < H; e += T) {
    const int k = keys[e];
    if (k >= 0) bp[e] = a.pos[k];
  }
  for (int e = threadIdx.x; e < HR; e += T) {
    const int r = rkeys[e];
    if (r >= 0) {
      rdone[e] = a.done[r];
      rviol[e] = a.violated[r] != 0;
    }
  }
  __syncthreads();
  for (int it = 0; it < iters; ++it) {
    for (int c = 0; c < ncolors; ++c) {
      if (rng[c].x == rng[c].y) continue;
      if (mc == c && rdone[sr] == 0) {
        float4 PA = bp[sa], PB = bp[sb];
        const float minSep = position_core(info.x & 0xff, info.y, m, lp, g, lc, rad, PA, PB);
        if (info.x & 0x100) bp[sa] = PA;
        if (info.x & 0x200) bp[sb] = PB;
        if (!(minSep >= -3.0f * kLinearSlop)) rviol[sr] = 1;
      }
      __syncthreads();
    }
    if (it + 1 < iters) {
      for (int e = threadIdx.x; e < HR; e += T)
        if (rkeys[e] >= 0 && rdone[e] == 0) {
          if (!rviol[e]) rdone[e] = it + 1;
          rviol[e] = 0;
        }
      __syncthreads();
    }
  }
  for (int e = threadIdx.x; e < H; e += T)
    if (dirty[e]) a.pos[keys[e]] = bp[e];
  for (int e = threadIdx.x; e < HR; e += T) {
    const int r = rkeys[e];
    if (r >= 0) {
      a.done[r] = rdone[e];
      a.violated[r] = rviol[e];
    }
  }
}

// Small islands' counts (at most cap constraints) for the bin scan; larger islands' constraints
// are counted into *big.
__global__ void __launch_bounds__(kBlock) k_island_small(int nb, const int* __restrict__ parent,
                                                         const int* __restrict__ isl_count, int cap,
                                                         int* __restrict__ small, int* __restrict__ big) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  int n = 0;
  if (b < nb && parent[b] == b) n = isl_count[b];
  if (b < nb) small[b] = n <= cap ? n : 0;
  if (n > cap) atomicAdd(big, n);
}

// (colour, bin) histogram and scatter of the constraints. A small island's constraints go to its
// bin (isl_start / cap); a large island's to the last group of their colour, whose few counters
// would see one global atomic per constraint: those are counted and ranked in LDS first and take
// one global atomic per colour and block, as k_color_hist and k_color_scatter do.
__device__ inline int constraint_group(int j, const int* __restrict__ state, const int* __restrict__ launch,
                                       const int* __restrict__ croot, const int* __restrict__ isl_count,
                                       const int* __restrict__ isl_start, int cap, int nbins) {
  const int L = launch[state[j] & 0xff], r = croot[j];
  return L * (nbins + 1) + (isl_count[r] > cap ? nbins : isl_start[r] / cap);
}

__global__ void __launch_bounds__(kBlock) k_color_bin_hist(int n, const int* __restrict__ state,
                                                           const int* __restrict__ launch, const int* __restrict__ croot,
                                                           const int* __restrict__ isl_count,
                                                           const int* __restrict__ isl_start, int cap, int nbins,
                                                           int* __restrict__ count) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = j < n && state[j] >= 0 ? constraint_group(j, state, launch, croot, isl_count, isl_start, cap, nbins) : -1;
  block_count_keys(g, count);
}

__global__ void __launch_bounds__(kBlock) k_color_bin_scatter(int n, const int* __restrict__ state,
                                                              const int* __restrict__ launch,
                                                              const int* __restrict__ croot,
                                                              const int* __restrict__ isl_count,
                                                              const int* __restrict__ isl_start, int cap, int nbins,
                                                              const int* __restrict__ off, int* __restrict__ cursor,
                                                              const int* __restrict__ list, int* __restrict__ order,
                                                              int* __restrict__ inv_order, int* __restrict__ pair_colour) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = j < n && state[j] >= 0 ? constraint_group(j, state, launch, croot, isl_count, isl_start, cap, nbins) : -1;
  const int pos = block_place_keys(g, off, cursor);
  if (g >= 0) {
    order[pos] = j;
    inv_order[j] = pos;
    pair_colour[list[j]] = g / (nbins + 1);
  }
}

__global__ void __launch_bounds__(kBlock) k_clear_two(int n, int* __restrict__ a, int* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    a[i] = 0;
    b[i] = 0;
  }
}

// per-colour launches of the large islands' constraints (the last group of each colour); the
// range is read from the device, the grid is sized for the whole colour
__global__ void __launch_bounds__(kBlock) k_solve_velocity_group(int c, const int* __restrict__ off, int ng,
                                                                 int block_solve, VelArgs a) {
  const int i = off[c * ng + ng - 1] + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < off[(c + 1) * ng]) solve_velocity_one(i, block_solve, a);
}
__global__ void __launch_bounds__(kBlock) k_solve_position_group(int c, const int* __restrict__ off, int ng,
                                                                 PosArgs a) {
  const int b0 = off[c * ng + ng - 1], b1 = off[(c + 1) * ng];
  if (b0 >= b1) return;
  const int i = b0 + blockIdx.x * blockDim.x + threadIdx.x;
  solve_position_one(i, i < b1, b0, a);
}

static int bin_threads_ok(int t) { return t == 64 || t == 128 || t == 256 || t == 512; }

static void launch_bins_velocity(World* w, long long items, int nbins, const int* off, int ncolors, int iters,
                                 int block_solve, const VelArgs& a) {
  const int ng = nbins + 1, K = B2G_KERNEL_SOLVE_VELOCITY;
  switch (w->bin_threads) {
    case 64:
      launch_timed(w, K, items, k_solve_bins_velocity<64>, dim3(nbins), dim3(64), off, ng, ncolors, iters, block_solve, a);
      break;
    case 128:
      launch_timed(w, K, items, k_solve_bins_velocity<128>, dim3(nbins), dim3(128), off, ng, ncolors, iters, block_solve, a);
      break;
    case 512:
      launch_timed(w, K, items, k_solve_bins_velocity<512>, dim3(nbins), dim3(512), off, ng, ncolors, iters, block_solve, a);
      break;
    default:
      launch_timed(w, K, items, k_solve_bins_velocity<256>, dim3(nbins), dim3(256), off, ng, ncolors, iters, block_solve, a);
      break;
  }
}
static void launch_bins_position(World* w, long long items, int nbins, const int* off, int ncolors, int iters,
                                 const PosArgs& a) {
  const int ng = nbins + 1, K = B2G_KERNEL_SOLVE_POSITION;
  switch (w->bin_threads) {
    case 64:
      launch_timed(w, K, items, k_solve_bins_position<64>, dim3(nbins), dim3(64), off, ng, ncolors, iters, a);
      break;
    case 128:
      launch_timed(w, K, items, k_solve_bins_position<128>, dim3(nbins), dim3(128), off, ng, ncolors, iters, a);
      break;
    case 512:
      launch_timed(w, K, items, k_solve_bins_position<512>, dim3(nbins), dim3(512), off, ng, ncolors, iters, a);
      break;
    default:
      launch_timed(w, K, items, k_solve_bins_position<256>, dim3(nbins), dim3(256), off, ng, ncolors, iters, a);
      break;
  }
}

// ---- graph-mode variants: colour ranges and counts read from device memory -----------------
// A captured graph replays with new colour offsets (copied before each launch) and a fixed
// grid; the loops stride over whatever the colour holds this step.
__global__ void __launch_bounds__(kBlock) k_solve_velocity_g(int c, const int* __restrict__ off, int block_solve,
                                                             VelArgs a) {
  const int b0 = off[c], b1 = off[c + 1];
  for (int i = b0 + blockIdx.x * blockDim.x + threadIdx.x; i < b1; i += gridDim.x * blockDim.x)
    solve_velocity_one(i, block_solve, a);
}
__global__ void __launch_bounds__(kBlock) k_solve_position_g(int c, const int* __restrict__ off, PosArgs a) {
  const int b0 = off[c], b1 = off[c + 1];
  // block-uniform trip count: whole waves take part in the island reduction
  for (int base = blockIdx.x * blockDim.x; base < b1 - b0; base += gridDim.x * blockDim.x) {
    const int i = b0 + base + threadIdx.x;
    solve_position_one(i, i < b1, b0, a);
  }
}
__global__ void __launch_bounds__(kBlock) k_store_impulses_g(const int* __restrict__ ncp, const int* __restrict__ kc,
                                                             const int4* __restrict__ info, const float4* __restrict__ lam,
                                                             float4* __restrict__ pt0, float4* __restrict__ pt1) {
  const int n = *ncp;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    store_one(i, kc, info, lam, pt0, pt1);
}

// everything a captured solve depends on: a cached graph is replayed only for the same key
struct SolveGraphKey {
  int ncolors, grid_col, grid_nc, nb, vel_iters, pos_iters, warm, block;
  float dt;
  const void* ptr[32];
  bool operator==(const SolveGraphKey& o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

// launches of the solve after the constraint set-up: warm start, velocity iterations, impulse
// store, position integration, position iterations, transforms (graph-mode kernels)
static void launch_solve_tail(World* w, const SolveGraphKey& k, const VelArgs& va, const PosArgs& pa) {
  SolverArrays& s = w->s;
  ContactArrays& c = w->c[w->cur];
  int* cnt = w->counters.p;
  hipStream_t st = w->stream;
  if (k.warm)
    hipLaunchKernelGGL(k_warm_start_bodies, dim3(blocks_for(k.nb, kBlock)), dim3(kBlock), 0, st, k.nb, w->b.type.p,
                       s.adj_off.p, s.adj.p, s.inv_order.p, va, IslandView{s.parent.p, nullptr});
  for (int it = 0; it < k.vel_iters; ++it)
    for (int col = 0; col < k.ncolors; ++col)
      hipLaunchKernelGGL(k_solve_velocity_g, dim3(k.grid_col), dim3(kBlock), 0, st, col, s.color_off.p, k.block, va);
  hipLaunchKernelGGL(k_store_impulses_g, dim3(k.grid_nc), dim3(kBlock), 0, st, cnt + 0, s.k_contact.p, s.k_info.p,
                     s.lam.p, c.pt0.p, c.pt1.p);
  const IslandView all_awake{s.parent.p, nullptr};
  hipLaunchKernelGGL(k_integrate_positions, dim3(blocks_for(k.nb, kBlock)), dim3(kBlock), 0, st, k.nb, w->b.type.p,
                     w->b.pos.p, w->b.vel.p, k.dt, all_awake);
  for (int it = 0; it < k.pos_iters; ++it) {
    for (int col = 0; col < k.ncolors; ++col)
      hipLaunchKernelGGL(k_solve_position_g, dim3(k.grid_col), dim3(kBlock), 0, st, col, s.color_off.p, pa);
    hipLaunchKernelGGL(k_position_iteration_end, dim3(blocks_for(k.nb, kBlock)), dim3(kBlock), 0, st, k.nb, it,
                       w->b.type.p, s.parent.p, s.isl_violated.p, s.isl_done.p, nullptr, nullptr, 0);
  }
  hipLaunchKernelGGL(k_finalize, dim3(blocks_for(k.nb, kBlock)), dim3(kBlock), 0, st, k.nb, w->b.type.p, w->b.pos.p,
                     w->b.mass.p, w->b.xf.p, all_awake);
}

// An executable graph is destroyed only once the stream has finished every launch of it: the
// host runs ahead of the device, so the last replay of a graph can still be queued or running,
// and destroying it then frees the runtime's copies of its kernel arguments under the launch.
void release_solve_graphs(World* w) {
  if (!w->graphs.empty()) (void)hipStreamSynchronize(w->stream);
  for (World::CachedGraph& g : w->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  w->graphs.clear();
}

// replays (capturing first if needed) the solve tail as a HIP graph
static int solve_tail_graph(World* w, const SolveGraphKey& k, const VelArgs& va, const PosArgs& pa) {
  const unsigned char* kb = reinterpret_cast<const unsigned char*>(&k);
  World::CachedGraph* hit = nullptr;
  for (World::CachedGraph& g : w->graphs)
    if (g.key.size() == sizeof(k) && std::memcmp(g.key.data(), kb, sizeof(k)) == 0) hit = &g;
  if (!hit) {
    if (w->graphs.size() >= 16) {  // evict the least recently used, once its launches are done
      size_t lru = 0;
      for (size_t i = 1; i < w->graphs.size(); ++i)
        if (w->graphs[i].last_use < w->graphs[lru].last_use) lru = i;
      B2G_HIP(hipStreamSynchronize(w->stream));
      (void)hipGraphExecDestroy(w->graphs[lru].exec);
      w->graphs.erase(w->graphs.begin() + lru);
    }
    hipGraph_t graph = nullptr;
    B2G_HIP(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
    launch_solve_tail(w, k, va, pa);
    B2G_HIP(hipStreamEndCapture(w->stream, &graph));
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      return B2G_ERR_DEVICE;
    }
    World::CachedGraph g;
    g.key.assign(kb, kb + sizeof(k));
    g.exec = exec;
    w->graphs.push_back(g);
    hit = &w->graphs.back();
  }
  hit->last_use = ++w->graph_clock;
  B2G_HIP(hipGraphLaunch(hit->exec, w->stream));
  return B2G_OK;
}

// ---- driver ---------------------------------------------------------------------------------------
static int reserve_solver(World* w, int nc) {
  SolverArrays& s = w->s;
  const size_t n = pair_capacity(w, nc);  // nc: the pairs, which bound the constraints
  int st;
  if ((st = s.touch_scan.reserve(pair_capacity(w, w->n_pairs) + 1)) || (st = s.touch_list.reserve(n)) ||
      (st = s.k_contact.reserve(n)) || (st = s.k_bodies.reserve(n)) || (st = s.k_info.reserve(n)) ||
      (st = s.vnf.reserve(n)) || (st = s.vm.reserve(n)) || (st = s.vr0.reserve(n)) || (st = s.vr1.reserve(n)) ||
      (st = s.vms.reserve(n)) || (st = s.vbias.reserve(n)) || (st = s.K.reserve(n)) || (st = s.NM.reserve(n)) ||
      (st = s.lam.reserve(n)) || (st = s.plp.reserve(n)) || (st = s.pgeo.reserve(n)) || (st = s.plc.reserve(n)) ||
      (st = s.prad.reserve(n)) || (st = s.adj.reserve(2 * n)) || (st = s.color_state.reserve(n)) ||
      (st = s.order.reserve(n)) || (st = s.inv_order.reserve(n)))
    return st;
  return B2G_OK;
}

int phase_solve(World* w, float dt, float dt_ratio, int vel_iters, int pos_iters) {
  const int nb = w->nb;
  if (nb == 0) return B2G_OK;
  SolverArrays& s = w->s;
  ContactArrays& c = w->c[w->cur];
  int st;
  int* cnt = w->counters.p;  // [0] touching, [1] uncoloured, [2] colour overflow, [3] unused
  // sleeping bodies: islands without an awake body are left out of the solve
  const bool sleepy = w->sleep_possible();
  // touching contacts: compacted on the device; the constraint arrays are sized for all pairs
  const int n_pairs = w->n_pairs;
  if ((st = reserve_solver(w, n_pairs))) return st;
  if ((st = s.deg.reserve(nb + 1)) || (st = s.adj_off.reserve(nb + 1)) || (st = s.adj_cursor.reserve(nb)) ||
      (st = s.parent.reserve(nb)) || (st = s.isl_violated.reserve(nb)) || (st = s.isl_done.reserve(nb)) ||
      (sleepy && ((st = s.isl_awake.reserve(nb)) || (st = s.isl_restless.reserve(nb)))) ||
      (st = s.color_off.reserve(3 * kMaxColors + 1)) ||
      (st = s.color_cursor.reserve(kMaxColors)) || (st = s.prev_colour.reserve(pair_capacity(w, n_pairs))))
    return st;
  const IntegrateArgs ia{w->b.type.p, w->b.pos.p, w->b.vel.p, w->b.pos0.p, w->b.props.p, w->b.mass.p, w->b.force.p,
                         w->gravity.x, w->gravity.y, dt};
  // batched worlds: each world's iterations in one workgroup (k_solve_worlds_*); otherwise small
  // islands packed into bins (k_solve_bins_*); neither with joints
  const bool want_world = w->nj == 0 && w->worlds_contiguous && w->n_worlds > 1 &&
                          (w->world_solve == 1 || (w->world_solve == -1 && w->n_worlds >= kWorldSolveMin));
  const bool want_bins = w->nj == 0 && (w->island_solve == 1 || (w->island_solve == -1 && !want_world));
  const int bin_cap = w->bin_threads / 2;  // islands of at most bin_cap constraints are binned
  if (want_bins && ((st = s.isl_count.reserve(nb)) || (st = s.isl_small.reserve(nb)) ||
                    (st = s.isl_start.reserve(nb)) || (st = s.croot.reserve(pair_capacity(w, n_pairs)))))
    return st;
  // per body: island roots, degrees, adjacency cursors, island counts and (no sleeping possible)
  // the velocity integration
  hipLaunchKernelGGL(k_solver_body_init, dim3(blocks_for(nb, kBlock)), dim3(kBlock), 0, w->stream, nb, s.parent.p,
                     s.deg.p, s.adj_cursor.p, int(!sleepy), ia, want_bins ? s.isl_count.p : nullptr);
  if ((st = exclusive_scan(w, s.touch_flag.p, s.touch_scan.p, n_pairs, cnt + 0))) return st;
  if (n_pairs == 0)
    hipLaunchKernelGGL(k_solver_counters_init, dim3(1), dim3(kBlock), 0, w->stream, cnt, s.color_cursor.p);
  const int gp = blocks_for(std::max(n_pairs, 1), kBlock);
  // Colour reuse (persisting contacts keep their colour) lets the colour count creep up as new
  // contacts take first-fit colours around kept ones; the step colours from scratch again once
  // the count exceeds the last from-scratch count by kColorSlack, and at least every
  // kColorRefresh steps (which keeps that count, the cap on kept colours, current).
  constexpr int kColorSlack = 2, kColorRefresh = 32;
  const bool reuse = w->color_reuse && w->scratch_colors > 0 && w->n_colors <= w->scratch_colors + kColorSlack &&
                     w->steps_since_scratch < kColorRefresh;
  // constraints (touching contacts): list, bodies, degrees, previous colours; then islands
  if (n_pairs > 0) {
    // colours of the previous step's contacts (the old set's solve wrote them)
    const int* old_colour = reuse ? w->c[1 - w->cur].colour.p : nullptr;
    hipLaunchKernelGGL(k_touching_setup, dim3(gp), dim3(kBlock), 0, w->stream, n_pairs, s.touch_flag.p,
                       s.touch_scan.p, s.touch_list.p, c.ab.p, w->f.body.p, w->b.type.p, s.k_bodies.p, s.deg.p,
                       c.old_index.p, old_colour, s.prev_colour.p, cnt, s.color_cursor.p);
    hipLaunchKernelGGL(k_union, dim3(gp), dim3(kBlock), 0, w->stream, cnt, s.k_bodies.p, w->b.type.p, s.parent.p);
  }
  launch_joint_union(w);
  hipLaunchKernelGGL(k_compress, dim3(blocks_for(nb, kBlock)), dim3(kBlock), 0, w->stream, nb, s.parent.p,
                     s.isl_violated.p, s.isl_done.p, sleepy ? s.isl_awake.p : nullptr, s.isl_restless.p);
  const IslandView iv{s.parent.p, sleepy ? s.isl_awake.p : nullptr};
  if (sleepy) {
    hipLaunchKernelGGL(k_island_awake, dim3(blocks_for(nb, kBlock)), dim3(kBlock), 0, w->stream, nb, w->b.type.p,
                       w->b.awake.p, s.parent.p, s.isl_awake.p);
    hipLaunchKernelGGL(k_integrate_velocities, dim3(blocks_for(nb, kBlock)), dim3(kBlock), 0, w->stream, nb, ia, iv,
                       w->b.awake.p);
  }
  int ncolors = 0, nc = 0, n_touching = 0;
  if (n_pairs > 0) {
    if ((st = exclusive_scan(w, s.deg.p, s.adj_off.p, nb + 1, nullptr))) return st;
    hipLaunchKernelGGL(k_fill_adjacency, dim3(gp), dim3(kBlock), 0, w->stream, cnt, s.k_bodies.p, w->b.type.p,
                       s.adj_off.p, s.adj_cursor.p, s.adj.p);
    // round 0: contacts that persist keep their colour; then Jones-Plassmann-Luby rounds for the
    // others, a first batch without waiting (rounds after the last constraint is coloured find
    // nothing to do), then more only if some remain
    hipLaunchKernelGGL(k_color_keep, dim3(gp), dim3(kBlock), 0, w->stream, cnt, s.k_bodies.p, w->b.type.p,
                       s.adj_off.p, s.adj.p, s.prev_colour.p, w->scratch_colors, iv, s.color_state.p,
                       want_bins ? s.isl_count.p : nullptr, s.croot.p);
    if (want_bins) {
      // the small islands' counts and their scan (bin of island r: isl_start[r] / bin_cap); the
      // large islands' constraint count comes back with the colouring's counters
      hipLaunchKernelGGL(k_island_small, dim3(blocks_for(nb, kBlock)), dim3(kBlock), 0, w->stream, nb, s.parent.p,
                         s.isl_count.p, bin_cap, s.isl_small.p, cnt + kBigIslandSlot);
      if ((st = exclusive_scan(w, s.isl_small.p, s.isl_start.p, nb, nullptr))) return st;
    }
    // first batch: from scratch the configured count; with reuse one more than the rounds the
    // last reusing step needed (the batch size does not change the colouring)
    int round = 1;
    for (;;) {
      int batch = round > 1 ? 4 : reuse ? std::max(2, std::min(w->color_rounds, w->reuse_rounds + 1)) : w->color_rounds;
      for (int r = 0; r < batch; ++r, ++round) {
        launch_timed(w, B2G_KERNEL_COLOR_ROUND, n_pairs, k_color_round, dim3(gp), dim3(kBlock), cnt, round,
                     s.k_bodies.p, w->b.type.p, s.adj_off.p, s.adj.p, s.color_state.p, cnt + 2);
      }
      // the histogram was cleared with the counters; a later batch recounts from zero
      if (round > batch + 1) {
        B2G_HIP(hipMemsetAsync(cnt + 4, 0, sizeof(int) * (kMaxColors + 1), w->stream));
        B2G_HIP(hipMemsetAsync(cnt + 1, 0, sizeof(int), w->stream));
      }
      hipLaunchKernelGGL(k_color_hist, dim3(gp), dim3(kBlock), 0, w->stream, cnt, s.color_state.p, cnt + 4, cnt + 1);
      B2G_HIP(hipMemcpyAsync(w->h_counters, cnt, (kBigIslandSlot + 1) * sizeof(int), hipMemcpyDeviceToHost,
                             w->stream));
      B2G_HIP(hipStreamSynchronize(w->stream));
      if ((st = kt_harvest(w))) return st;
      if (w->h_counters[1] == 0) break;
      if (round > 4 * n_pairs + 64) {
        set_error("colouring did not converge");
        return B2G_ERR_DEVICE;
      }
    }
    w->color_rounds_used = round - 1;
    if (reuse) w->reuse_rounds = w->h_counters[4 + kMaxColors];
    // constraints to solve: the coloured ones (touching contacts of awake islands)
    n_touching = w->h_counters[0];
    for (int i = 0; i < kMaxColors; ++i) nc += w->h_counters[4 + i];
  }
  // the highest colour may hold constraints that share a body (a body with more contacts than
  // the colour masks hold): it is solved serially, so the island bins, the per-world solve and
  // the solve graph, which solve a colour's constraints in parallel, are not used
  const bool serial = nc > 0 && w->h_counters[4 + kMaxColors - 1] > 0;
  if (serial && w->tape.active) {
    set_error("the adjoint tape needs every colour conflict-free (a body has more than 127 contacts)");
    return B2G_ERR_UNSUPPORTED;
  }
  w->n_touch = nc;
  const int n_big = want_bins && nc > 0 ? w->h_counters[kBigIslandSlot] : 0;
  const int n_small = want_bins ? nc - n_big : 0;
  // auto: bins only where they take at least half of the constraints (the per-colour launches the
  // large islands need take the few small ones at no extra launch)
  const bool bins = !serial && n_small > 0 && (w->island_solve == 1 || n_small >= n_big);
  const int nbins = bins ? (n_small + bin_cap - 1) / bin_cap : 0;
  w->n_bins = nbins;
  const bool per_world = nc > 0 && want_world && !want_bins && !serial;
  if (nc > 0) {
    // One launch per non-empty colour, in colour-id order (kept colours can leave gaps).
    // Device copy: launch offsets (kMaxColors + 1, padded with nc), then per colour id its base
    // offset and its launch index.
    const int* hc = w->h_counters + 4;
    // pinned (the copy reads it asynchronously; the next step writes it only after its colouring's
    // read-back, when this copy has long finished)
    static_assert(kPinnedColourOffsets + 3 * kMaxColors + 1 <= kPinnedInts, "pinned colour offsets");
    int* offs = w->h_counters + kPinnedColourOffsets;
    std::fill(offs, offs + 3 * kMaxColors + 1, 0);
    int* base = offs + kMaxColors + 1;
    int* launch = base + kMaxColors;
    w->h_color_off.assign(1, 0);
    for (int i = 0; i < kMaxColors; ++i) {
      base[i] = w->h_color_off.back();
      launch[i] = ncolors;
      if (hc[i] == 0) continue;
      ++ncolors;
      w->h_color_off.push_back(w->h_color_off.back() + hc[i]);
    }
    for (int i = 0; i <= kMaxColors; ++i) offs[i] = i <= ncolors ? w->h_color_off[i] : nc;
    B2G_HIP(hipMemcpyAsync(s.color_off.p, offs, sizeof(int) * (3 * kMaxColors + 1), hipMemcpyHostToDevice,
                           w->stream));
    if (bins) {
      // constraints grouped by bin inside each colour, the large islands' last (see k_color_bin_hist)
      const int ng = nbins + 1, ncw = ncolors * ng;
      if ((st = s.cw_count.reserve(ncw)) || (st = s.cw_cursor.reserve(ncw)) || (st = s.cw_off.reserve(ncw + 1)))
        return st;
      hipLaunchKernelGGL(k_clear_two, dim3(blocks_for(ncw, kBlock)), dim3(kBlock), 0, w->stream, ncw, s.cw_count.p,
                         s.cw_cursor.p);
      const int* launch = s.color_off.p + kMaxColors + 1 + kMaxColors;
      hipLaunchKernelGGL(k_color_bin_hist, dim3(blocks_for(n_touching, kBlock)), dim3(kBlock), 0, w->stream,
                         n_touching, s.color_state.p, launch, s.croot.p, s.isl_count.p, s.isl_start.p, bin_cap, nbins,
                         s.cw_count.p);
      if ((st = exclusive_scan(w, s.cw_count.p, s.cw_off.p, ncw, s.cw_off.p + ncw))) return st;
      hipLaunchKernelGGL(k_color_bin_scatter, dim3(blocks_for(n_touching, kBlock)), dim3(kBlock), 0, w->stream,
                         n_touching, s.color_state.p, launch, s.croot.p, s.isl_count.p, s.isl_start.p, bin_cap, nbins,
                         s.cw_off.p, s.cw_cursor.p, s.touch_list.p, s.order.p, s.inv_order.p, c.colour.p);
    } else if (per_world) {
      // constraints grouped by world inside each colour (see k_color_world_hist)
      const int nw = w->n_worlds, ncw = ncolors * nw;
      if ((st = s.cw_count.reserve(ncw)) || (st = s.cw_cursor.reserve(ncw)) || (st = s.cw_off.reserve(ncw + 1)))
        return st;
      hipLaunchKernelGGL(k_clear_ints, dim3(blocks_for(ncw, kBlock)), dim3(kBlock), 0, w->stream, ncw, s.cw_count.p);
      hipLaunchKernelGGL(k_clear_ints, dim3(blocks_for(ncw, kBlock)), dim3(kBlock), 0, w->stream, ncw, s.cw_cursor.p);
      const int* launch = s.color_off.p + kMaxColors + 1 + kMaxColors;
      hipLaunchKernelGGL(k_color_world_hist, dim3(blocks_for(n_touching, kBlock)), dim3(kBlock), 0, w->stream,
                         n_touching, s.color_state.p, launch, s.touch_list.p, c.ab.p, w->f.world.p, nw, s.cw_count.p);
      if ((st = exclusive_scan(w, s.cw_count.p, s.cw_off.p, ncw, s.cw_off.p + ncw))) return st;
      hipLaunchKernelGGL(k_color_world_scatter, dim3(blocks_for(n_touching, kBlock)), dim3(kBlock), 0, w->stream,
                         n_touching, s.color_state.p, launch, s.touch_list.p, c.ab.p, w->f.world.p, nw, s.cw_off.p,
                         s.cw_cursor.p, s.order.p, s.inv_order.p, c.colour.p);
    } else {
      hipLaunchKernelGGL(k_color_scatter, dim3(blocks_for(n_touching, kBlock)), dim3(kBlock), 0, w->stream, n_touching,
                         s.color_state.p, s.color_off.p + kMaxColors + 1, s.color_cursor.p, s.order.p, s.inv_order.p,
                         s.touch_list.p, c.colour.p);
    }
    InitArgs ia;
    ia.n = nc;
    ia.order = s.order.p;
    ia.list = s.touch_list.p;
    ia.ab = c.ab.p;
    ia.fbody = w->f.body.p;
    ia.fshape = w->f.shape.p;
    ia.fmat = w->f.mat.p;
    ia.btype = w->b.type.p;
    ia.parent = s.parent.p;
    ia.bpos = w->b.pos.p;
    ia.bvel = w->b.vel.p;
    ia.bmass = w->b.mass.p;
    ia.hdr = c.hdr.p;
    ia.geo = c.geo.p;
    ia.pt0 = c.pt0.p;
    ia.pt1 = c.pt1.p;
    ia.dt_ratio = dt_ratio;
    ia.warm = w->warm_starting;
    ia.block_solve = w->block_solve;
    ia.k_contact = s.k_contact.p;
    ia.k_bodies = s.k_bodies.p;  // overwritten in colour order (the unsorted copy is no longer needed)
    ia.k_info = s.k_info.p;
    ia.vnf = s.vnf.p; ia.vm = s.vm.p; ia.vr0 = s.vr0.p; ia.vr1 = s.vr1.p; ia.vms = s.vms.p; ia.vbias = s.vbias.p;
    ia.K = s.K.p; ia.NM = s.NM.p; ia.lam = s.lam.p; ia.plp = s.plp.p; ia.pgeo = s.pgeo.p; ia.plc = s.plc.p;
    ia.prad = s.prad.p;
    // k_bodies is both read (colouring, done) and rewritten here in sorted order: the init
    // kernel recomputes bodies from the contact, so the aliasing is safe.
    launch_timed(w, B2G_KERNEL_INIT_CONSTRAINTS, nc, k_init_constraints, dim3(blocks_for(nc, kBlock)), dim3(kBlock),
                 ia);
  }
  w->n_colors = ncolors;
  if (!reuse && ncolors > 0) {
    w->scratch_colors = ncolors;
    w->steps_since_scratch = 0;
  } else {
    ++w->steps_since_scratch;
  }
  VelArgs va{s.k_bodies.p, s.k_info.p, s.vnf.p, s.vm.p, s.vr0.p, s.vr1.p, s.vms.p, s.vbias.p, s.K.p, s.NM.p, s.lam.p,
             w->b.vel.p};
  PosArgs pa{s.k_bodies.p, s.k_info.p, s.vm.p, s.plp.p, s.pgeo.p, s.plc.p, s.prad.p, w->b.pos.p, s.isl_violated.p,
             s.isl_done.p};
  const std::vector<int>& co = w->h_color_off;
  // (graph nodes cannot carry the timing events of a timed solve kernel)
  const bool timed_tail = w->kt.kernel == B2G_KERNEL_WARM_START || w->kt.kernel == B2G_KERNEL_SOLVE_VELOCITY ||
                          w->kt.kernel == B2G_KERNEL_SOLVE_POSITION;
  if (nc > 0 && w->solver_graphs && !timed_tail && !sleepy && w->nj == 0 && !per_world && !bins && !serial) {
    // one graph launch replaces ~(1 + vel_iters + pos_iters) x colours launches
    int maxc = 0;
    for (int col = 0; col < ncolors; ++col) maxc = std::max(maxc, co[col + 1] - co[col]);
    auto bucket = [](int n) {
      int b = 256;
      while (b < n) b <<= 1;
      return b;
    };
    SolveGraphKey k;
    std::memset(&k, 0, sizeof(k));
    k.ncolors = ncolors;
    k.grid_col = blocks_for(bucket(maxc), kBlock);
    k.grid_nc = blocks_for(bucket(nc), kBlock);
    k.nb = nb;
    k.vel_iters = vel_iters;
    k.pos_iters = pos_iters;
    k.warm = w->warm_starting;
    k.block = w->block_solve;
    k.dt = dt;
    const void* ptrs[] = {va.bodies, va.info, va.vnf, va.vm, va.vr0, va.vr1, va.vms, va.vbias, va.K, va.NM, va.lam,
                          va.vel, pa.plp, pa.pgeo, pa.plc, pa.prad, pa.pos, pa.violated, pa.done, s.color_off.p, cnt,
                          c.pt0.p, c.pt1.p, s.k_contact.p, w->b.type.p, s.parent.p, w->b.mass.p, w->b.xf.p,
                          s.adj_off.p, s.adj.p, s.inv_order.p};
    static_assert(sizeof(ptrs) / sizeof(ptrs[0]) <= 32, "graph key pointers");
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) k.ptr[i] = ptrs[i];
    if ((st = solve_tail_graph(w, k, va, pa))) return st;
    w->pos_iters_run = pos_iters;
    return B2G_OK;
  }
  if (nc > 0 && w->warm_starting)
    launch_timed(w, B2G_KERNEL_WARM_START, nc, k_warm_start_bodies, dim3(blocks_for(nb, kBlock)), dim3(kBlock), nb,
                 w->b.type.p, s.adj_off.p, s.adj.p, s.inv_order.p, va, iv);
  // joints (b2g_joints.hip): initialised after the contacts' warm start, solved before the
  // contacts in each velocity iteration and after them in each position iteration
  const JointArgs ja{w->j.order.p, w->j.bodies.p, w->j.type.p, w->j.d0.p,   w->j.d1.p,   w->j.d2.p,
                     w->j.d3.p,    w->j.imp.p,    w->j.imp2.p, w->j.r.p,    w->j.k.p,    w->j.k2.p,
                     w->j.k3.p,    w->b.type.p,   w->b.pos.p,  w->b.vel.p,  w->b.mass.p, iv,
                     s.isl_violated.p, s.isl_done.p, dt, 1.0f / dt, dt_ratio, int(w->warm_starting)};
  if (w->nj > 0) launch_joint_colours(w, 0, ja);
  if (per_world) {
    launch_worlds_velocity(w, static_cast<long long>(nc) * vel_iters, s.cw_off.p, ncolors, vel_iters,
                           int(w->block_solve), va);
    hipLaunchKernelGGL(k_store_and_integrate, dim3(blocks_for(std::max(nc, nb), kBlock)), dim3(kBlock), 0, w->stream,
                       nc, s.k_contact.p, s.k_info.p, s.lam.p, c.pt0.p, c.pt1.p, nb, w->b.type.p, w->b.pos.p,
                       w->b.vel.p, dt, iv);
    const IterationEndArgs ends{-1, s.parent.p, s.isl_violated.p, s.isl_done.p, iv.isl_awake};
    if (pos_iters > 0)
      launch_worlds_position(w, static_cast<long long>(nc) * pos_iters, s.cw_off.p, ncolors, pos_iters, pa, ends);
    w->pos_iters_run = pos_iters;
  } else {
  // Per-colour launches: of all constraints, or with island bins of the large islands' (the last
  // of the ng groups of every colour), the bins' kernels solving the rest; the two sets share no
  // body, so they may run in either order.
  const int ng = bins ? nbins + 1 : 1;
  const int* goff = bins ? s.cw_off.p : s.color_off.p;
  const bool per_colour = !bins || n_big > 0;
  // colours [tail, last): the small last colours, one launch per iteration (when at least two);
  // [last, ncolors): the serial overflow colour; co holds ncolors + 1 offsets only when there are
  // constraints
  const int last = serial ? ncolors - 1 : ncolors;
  int tail = last;
  if (nc > 0) {
    while (tail > 0 && co[tail] - co[tail - 1] <= kTailMax) --tail;
    if (last - tail < 2) tail = last;
  }
  const int tail_items = nc > 0 ? co[last] - co[tail] : 0;
  if (bins && vel_iters > 0)
    launch_bins_velocity(w, static_cast<long long>(n_small) * vel_iters, nbins, s.cw_off.p, ncolors, vel_iters,
                         int(w->block_solve), va);
  for (int it = 0; it < vel_iters && per_colour && (nc > 0 || w->nj > 0); ++it) {
    if (w->nj > 0) launch_joint_colours(w, 1, ja);
    for (int col = 0; col < tail; ++col) {
      int b0 = co[col], b1 = co[col + 1];
      if (bins)
        launch_timed(w, B2G_KERNEL_SOLVE_VELOCITY, b1 - b0, k_solve_velocity_group, dim3(blocks_for(b1 - b0, kBlock)),
                     dim3(kBlock), col, goff, ng, int(w->block_solve), va);
      else
        launch_timed(w, B2G_KERNEL_SOLVE_VELOCITY, b1 - b0, k_solve_velocity, dim3(blocks_for(b1 - b0, kBlock)),
                     dim3(kBlock), b0, b1, int(w->block_solve), va);
    }
    if (tail < last)
      launch_timed(w, B2G_KERNEL_SOLVE_VELOCITY, tail_items, k_solve_velocity_tail, dim3(1), dim3(kTailBlock), tail,
                   last, goff, ng, int(w->block_solve), va);
    if (serial)
      hipLaunchKernelGGL(k_solve_velocity_serial, dim3(1), dim3(64), 0, w->stream, co[last], co[ncolors],
                         int(w->block_solve), va);
  }
  hipLaunchKernelGGL(k_store_and_integrate, dim3(blocks_for(std::max(nc, nb), kBlock)), dim3(kBlock), 0, w->stream, nc,
                     s.k_contact.p, s.k_info.p, s.lam.p, c.pt0.p, c.pt1.p, nb, w->b.type.p, w->b.pos.p, w->b.vel.p, dt,
                     iv);
  if (bins && pos_iters > 0)
    launch_bins_position(w, static_cast<long long>(n_small) * pos_iters, nbins, s.cw_off.p, ncolors, pos_iters, pa);
  w->pos_iters_run = 0;
  for (int it = 0; it < pos_iters; ++it) {
    if (per_colour) {
      for (int col = 0; col < tail; ++col) {
        int b0 = co[col], b1 = co[col + 1];
        if (bins)
          launch_timed(w, B2G_KERNEL_SOLVE_POSITION, b1 - b0, k_solve_position_group,
                       dim3(blocks_for(b1 - b0, kBlock)), dim3(kBlock), col, goff, ng, pa);
        else
          launch_timed(w, B2G_KERNEL_SOLVE_POSITION, b1 - b0, k_solve_position, dim3(blocks_for(b1 - b0, kBlock)),
                       dim3(kBlock), b0, b1, pa);
      }
      if (tail < last)
        launch_timed(w, B2G_KERNEL_SOLVE_POSITION, tail_items, k_solve_position_tail, dim3(1), dim3(kTailBlock), tail,
                     last, goff, ng, pa);
      if (serial)
        hipLaunchKernelGGL(k_solve_position_serial, dim3(1), dim3(64), 0, w->stream, co[last], co[ncolors], pa);
    }
    if (w->nj > 0) launch_joint_colours(w, 2, ja);
    // (the last iteration's end runs with the copy back below; the binned islands' ends run in
    // their kernel)
    if (it + 1 < pos_iters && per_colour)
      hipLaunchKernelGGL(k_position_iteration_end, dim3(blocks_for(nb, kBlock)), dim3(kBlock), 0, w->stream, nb, it,
                         w->b.type.p, s.parent.p, s.isl_violated.p, s.isl_done.p, iv.isl_awake,
                         bins ? s.isl_count.p : nullptr, bin_cap);
    ++w->pos_iters_run;
  }
  }
  const IterationEndArgs last_end{pos_iters - 1, s.parent.p, s.isl_violated.p, s.isl_done.p, iv.isl_awake};
  hipLaunchKernelGGL(k_finalize_end, dim3(blocks_for(nb, kBlock)), dim3(kBlock), 0, w->stream, nb, w->b.type.p,
                     w->b.pos.p, w->b.mass.p, w->b.xf.p, iv, last_end);
  if (w->allow_sleep) {
    hipLaunchKernelGGL(k_sleep_time, dim3(blocks_for(nb, kBlock)), dim3(kBlock), 0, w->stream, nb, w->b.type.p,
                       w->b.auto_sleep.p, w->b.vel.p, w->b.sleep_time.p, iv, s.isl_restless.p, dt);
    hipLaunchKernelGGL(k_sleep_apply, dim3(blocks_for(nb, kBlock)), dim3(kBlock), 0, w->stream, nb, w->b.type.p, iv,
                       s.isl_restless.p, s.isl_done.p, w->b.awake.p, w->b.sleep_time.p, w->b.vel.p, w->b.force.p);
  }
  B2G_HIP(hipGetLastError());
  return B2G_OK;
}

}  // namespace b2g
