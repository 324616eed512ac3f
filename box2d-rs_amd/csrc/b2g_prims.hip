// b2g primitives: device-wide exclusive scan and LSD radix sort of (key, value) pairs.
//
// Both are written for 64-wide wavefronts: block-level scans use wave prefix sums built on
// DPP-free shuffles plus one LDS exchange per wave, and the radix sort ranks keys within a
// wave by matching digits with __ballot (wave multi-split), so ranks are stable without
// per-key LDS atomics.
#include "b2g_internal.h"

namespace b2g {

static constexpr int kScanBlock = 256;
static constexpr int kScanItems = 8;   // items per thread
static constexpr int kScanTile = kScanBlock * kScanItems;

// inclusive prefix sum within a 64-lane wave
__device__ inline int wave_inclusive(int x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// exclusive block scan of one value per thread; returns the block total in *total
__device__ inline int block_exclusive(int x, int* total) {
  __shared__ int wsum[kScanBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int inc = wave_inclusive(x);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanBlock / 64; ++w) {
    int s = wsum[w];
    if (w < wid) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + inc - x;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const int* __restrict__ in, int n, int* __restrict__ sums) {
  const int base = blockIdx.x * kScanTile;
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int idx = base + i * kScanBlock + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  int tot;
  block_exclusive(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_apply(const int* __restrict__ in, int* __restrict__ out, int n,
                                                          const int* __restrict__ sums, int nblocks,
                                                          int* __restrict__ total) {
  // the block's offset: the sum of the block sums before it (at most a few hundred), summed by
  // the block itself instead of in a separate launch
  __shared__ int wpre[kScanBlock / 64];
  int pre = 0;
  for (int i = threadIdx.x; i < blockIdx.x; i += kScanBlock) pre += sums[i];
  pre = __shfl(wave_inclusive(pre), 63, 64);
  if ((threadIdx.x & 63) == 0) wpre[threadIdx.x >> 6] = pre;
  __syncthreads();
  pre = 0;
#pragma unroll
  for (int w = 0; w < kScanBlock / 64; ++w) pre += wpre[w];
  // each thread owns kScanItems consecutive items so the block covers a contiguous tile
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int idx = base + i;
    v[i] = idx < n ? in[idx] : 0;
    s += v[i];
  }
  int tot;
  int ex = block_exclusive(s, &tot) + pre;
  if (total && blockIdx.x == nblocks - 1 && threadIdx.x == 0) *total = pre + tot;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int idx = base + i;
    if (idx < n) out[idx] = ex;
    ex += v[i];
  }
}

// Single-pass scan with decoupled look-back: each block takes the next tile in launch order
// from a ticket counter, scans it, publishes its sum (flag A), sums the published values of the
// tiles before it, walking back until one that carries its inclusive prefix (flag P), then
// publishes its own prefix and writes its outputs. A tile only waits for tiles with smaller
// tickets, which are already running and publish their sums before they look back themselves,
// so the wait always ends. Published words carry the call's epoch in their top bits, so the
// states never need clearing between calls. The spin is bounded: past the bound the block sets
// the error flag (ctl[1], read back with the broadphase's counters) instead of hanging.
static constexpr unsigned long long kScanFlagA = 1ull << 32, kScanFlagP = 2ull << 32;
static constexpr int kScanSpinBound = 1 << 22;
__global__ void __launch_bounds__(kScanBlock) k_scan_lookback(const int* in, int* out, int n, int* __restrict__ total,
                                                             unsigned long long* __restrict__ state,
                                                             int* __restrict__ ctl, unsigned epoch, int ntiles,
                                                             int* __restrict__ total2) {
  __shared__ int tile_s, prefix_s;
  if (threadIdx.x == 0) tile_s = atomicAdd(&ctl[0], 1);
  __syncthreads();
  const int tile = tile_s;
  const int base = tile * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const int idx = base + i;
    v[i] = idx < n ? in[idx] : 0;
    sum += v[i];
  }
  int agg;
  int ex = block_exclusive(sum, &agg);
  // the first wave publishes the tile's sum and looks back 64 tiles at a time: lane l reads tile
  // (j - l); the window is summed up to the nearest tile that holds its inclusive prefix (or all
  // 64 when none does and all are published), and read again while one of those is unpublished
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    const unsigned long long tag = static_cast<unsigned long long>(epoch) << 34;
    int prefix = 0;
    if (tile > 0) {
      if (lane == 0)
        __hip_atomic_store(&state[tile], tag | kScanFlagA | static_cast<unsigned>(agg), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      int j = tile - 1, spins = 0;
      for (;;) {
        const int idx = j - lane;
        unsigned long long word = 0;
        bool valid = true, is_p = true;  // before tile 0: an empty prefix
        if (idx >= 0) {
          word = __hip_atomic_load(&state[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          valid = (word >> 34) == epoch && (word & (3ull << 32)) != 0;
          is_p = valid && (word & (3ull << 32)) == kScanFlagP;
        }
        const unsigned long long pmask = __ballot(is_p), vmask = __ballot(valid);
        const int first_p = pmask ? __ffsll(static_cast<long long>(pmask)) - 1 : 64;
        const unsigned long long need = first_p >= 63 ? ~0ull : ((1ull << (first_p + 1)) - 1);
        if ((vmask & need) != need) {
          if (++spins > kScanSpinBound) {
            if (lane == 0) ctl[1] = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        int val = (lane <= first_p && idx >= 0) ? static_cast<int>(static_cast<unsigned>(word)) : 0;
        for (int d = 32; d >= 1; d >>= 1) val += __shfl_xor(val, d, 64);
        prefix += val;
        if (first_p < 64) break;
        j -= 64;
      }
    }
    if (lane == 0) {
      __hip_atomic_store(&state[tile], tag | kScanFlagP | static_cast<unsigned>(prefix + agg), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      prefix_s = prefix;
      if (tile == ntiles - 1) {
        if (total) *total = prefix + agg;
        if (total2) *total2 = prefix + agg;
        ctl[0] = 0;  // every ticket is taken: ready for the next call
      }
    }
  }
  __syncthreads();
  ex += prefix_s;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const int idx = base + i;
    if (idx < n) out[idx] = ex;
    ex += v[i];
  }
}

// Scan of a short array by one block in one launch: thread t owns a contiguous chunk, sums it,
// the block scans the chunk sums, and each thread writes its chunk's prefixes (reading each
// element before writing it, so in and out may alias).
static constexpr int kOneBlock = 1024;
static constexpr int kOneBlockMax = 4 * kOneBlock;  // longer arrays: the three-launch scan is faster
__global__ void __launch_bounds__(kOneBlock) k_scan_one_block(const int* in, int* out, int n, int* __restrict__ total,
                                                               int* __restrict__ total2) {
  __shared__ int wsum[kOneBlock / 64];
  const int chunk = (n + kOneBlock - 1) / kOneBlock;
  const int b0 = min(n, int(threadIdx.x) * chunk), b1 = min(n, b0 + chunk);
  int s = 0;
  for (int i = b0; i < b1; ++i) s += in[i];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int inc = wave_inclusive(s);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int base = 0;
  for (int v = 0; v < wid; ++v) base += wsum[v];
  int ex = base + inc - s;
  if (threadIdx.x == kOneBlock - 1 && total) *total = ex + s;
  if (threadIdx.x == kOneBlock - 1 && total2) *total2 = ex + s;
  for (int i = b0; i < b1; ++i) {
    const int v = in[i];
    out[i] = ex;
    ex += v;
  }
}

// out[i] = sum of in[0..i); *total_dev and *total2 (if given) = sum of all. in and out may alias.
int exclusive_scan(World* w, const int* in, int* out, int n, int* total_dev, int* total2) {
  if (n <= 0) {
    if (total_dev) B2G_HIP(hipMemsetAsync(total_dev, 0, sizeof(int), w->stream));
    if (total2) B2G_HIP(hipMemsetAsync(total2, 0, sizeof(int), w->stream));
    return B2G_OK;
  }
  if (n <= kOneBlockMax) {
    hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(kOneBlock), 0, w->stream, in, out, n, total_dev, total2);
    B2G_HIP(hipGetLastError());
    return B2G_OK;
  }
  const int nblocks = blocks_for(n, kScanTile);
  BroadphaseArrays& bp = w->bp;
  int st;
  // the tile states are cleared when (re)allocated and when the epoch wraps; epoch 0 is never used
  const bool fresh = bp.scan_state.cap < size_t(nblocks);
  if ((st = bp.scan_state.reserve(nblocks)) || (st = bp.scan_ctl.reserve(2))) return st;
  if (!bp.scan_ctl_ready) {
    B2G_HIP(hipMemsetAsync(bp.scan_ctl.p, 0, 2 * sizeof(int), w->stream));
    bp.scan_ctl_ready = true;
  }
  if (fresh || bp.scan_epoch + 1 >= (1u << 30)) {
    B2G_HIP(hipMemsetAsync(bp.scan_state.p, 0, sizeof(unsigned long long) * bp.scan_state.cap, w->stream));
    bp.scan_epoch = 0;
  }
  ++bp.scan_epoch;
  hipLaunchKernelGGL(k_scan_lookback, dim3(nblocks), dim3(kScanBlock), 0, w->stream, in, out, n, total_dev,
                     bp.scan_state.p, bp.scan_ctl.p, bp.scan_epoch, nblocks, total2);
  B2G_HIP(hipGetLastError());
  return B2G_OK;
}

// ---- radix sort ---------------------------------------------------------------------------
static constexpr int kSortBlock = 256;
static constexpr int kSortItems = 4;
static constexpr int kSortTile = kSortBlock * kSortItems;
static constexpr int kRadix = 256;

// histogram of one 8-bit digit per tile; hist is digit-major: hist[d * nblocks + block]
__global__ void __launch_bounds__(kSortBlock) k_radix_hist(const unsigned* __restrict__ keys, int n, int shift,
                                                           unsigned* __restrict__ hist, int nblocks) {
  __shared__ unsigned h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * kSortTile;
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    int idx = base + i * kSortBlock + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 0xffu], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// stable scatter: the tile is processed as kSortItems sub-tiles of kSortBlock keys in index
// order; within a sub-tile each wave ranks its keys by digit with ballots. kFromHist: offs holds
// the per-tile histograms themselves (few tiles) and the block computes its digits' offsets
// (the keys of smaller digits plus the digit's keys in earlier tiles) instead of reading a scan.
static constexpr int kRadixFuseMax = 128;  // tiles up to which the scatter sums the histograms itself
static_assert(kSortBlock == kRadix && kSortBlock == kScanBlock, "one thread per digit, block scan of 256");
template <bool kFromHist>
__global__ void __launch_bounds__(kSortBlock) k_radix_scatter(const unsigned* __restrict__ keys,
                                                              const unsigned* __restrict__ vals, int n, int shift,
                                                              const unsigned* __restrict__ offs, int nblocks,
                                                              unsigned* __restrict__ okeys, unsigned* __restrict__ ovals) {
  __shared__ unsigned running[kRadix];
  __shared__ unsigned wcount[kSortBlock / 64][kRadix];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if constexpr (kFromHist) {
    const unsigned* row = offs + threadIdx.x * nblocks;
    const int me = static_cast<int>(blockIdx.x);
    unsigned before = 0, total = 0;
    if ((nblocks & 3) == 0) {
      // rows start 16-byte aligned: read them four tiles at a time, all loads in flight together
      const uint4* row4 = reinterpret_cast<const uint4*>(row);
#pragma unroll 8
      for (int q = 0; q < nblocks / 4; ++q) {
        const uint4 h = row4[q];
        const int t = 4 * q;
        total += h.x + h.y + h.z + h.w;
        before += (t < me ? h.x : 0u) + (t + 1 < me ? h.y : 0u) + (t + 2 < me ? h.z : 0u) + (t + 3 < me ? h.w : 0u);
      }
    } else {
#pragma unroll 8
      for (int t = 0; t < nblocks; ++t) {
        const unsigned h = row[t];
        total += h;
        before += t < me ? h : 0u;
      }
    }
    int all;
    running[threadIdx.x] = static_cast<unsigned>(block_exclusive(static_cast<int>(total), &all)) + before;
  } else {
    running[threadIdx.x] = offs[threadIdx.x * nblocks + blockIdx.x];
  }
  const int base = blockIdx.x * kSortTile;
  for (int sub = 0; sub < kSortItems; ++sub) {
    for (int w = 0; w < kSortBlock / 64; ++w) wcount[w][threadIdx.x] = 0;
    __syncthreads();
    const int idx = base + sub * kSortBlock + threadIdx.x;
    const bool valid = idx < n;
    unsigned key = valid ? keys[idx] : 0u;
    unsigned val = valid ? vals[idx] : 0u;
    unsigned d = (key >> shift) & 0xffu;
    // lanes of this wave holding the same digit
    unsigned long long same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      unsigned long long bb = __ballot((d >> b) & 1u);
      same &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned rank = __popcll(same & below);
    bool leader = valid && (same & below) == 0ull;
    if (leader) wcount[wid][d] = __popcll(same);
    __syncthreads();
    if (valid) {
      unsigned pre = running[d];
      for (int w = 0; w < wid; ++w) pre += wcount[w][d];
      okeys[pre + rank] = key;
      ovals[pre + rank] = val;
    }
    __syncthreads();
    unsigned add = 0;
    for (int w = 0; w < kSortBlock / 64; ++w) add += wcount[w][threadIdx.x];
    running[threadIdx.x] += add;
    __syncthreads();
  }
}

// Sorts n (key, value) pairs by the low `bits` bits of the key, stable. The result ends in
// keys/vals (the alternate buffers are scratch).
// Sorted (keys, vals) end in the given arrays; when those are the world's bp.keys / bp.vals and an
// odd number of passes leaves the result in the alternate buffers, the buffers are swapped
// instead of copied (callers read bp.keys.p / bp.vals.p again after the call).
int radix_sort_pairs(World* w, unsigned* keys, unsigned* vals, int n, int bits) {
  if (n <= 1) return B2G_OK;
  BroadphaseArrays& bp = w->bp;
  int st;
  if ((st = bp.keys_alt.reserve(n))) return st;
  if ((st = bp.vals_alt.reserve(n))) return st;
  int nblocks = blocks_for(n, kSortTile);
  if ((st = bp.hist.reserve(size_t(kRadix) * nblocks))) return st;
  unsigned *k0 = keys, *v0 = vals, *k1 = bp.keys_alt.p, *v1 = bp.vals_alt.p;
  int passes = (bits + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    int shift = 8 * p;
    hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kSortBlock), 0, w->stream, k0, n, shift, bp.hist.p, nblocks);
    if (nblocks <= kRadixFuseMax) {
      hipLaunchKernelGGL(k_radix_scatter<true>, dim3(nblocks), dim3(kSortBlock), 0, w->stream, k0, v0, n, shift,
                         bp.hist.p, nblocks, k1, v1);
    } else {
      if ((st = exclusive_scan(w, reinterpret_cast<int*>(bp.hist.p), reinterpret_cast<int*>(bp.hist.p),
                               kRadix * nblocks, nullptr)))
        return st;
      hipLaunchKernelGGL(k_radix_scatter<false>, dim3(nblocks), dim3(kSortBlock), 0, w->stream, k0, v0, n, shift,
                         bp.hist.p, nblocks, k1, v1);
    }
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  if (k0 != keys) {
    if (keys == bp.keys.p && vals == bp.vals.p) {
      // the result is in the alternate buffers: they become the world's keys and values
      std::swap(bp.keys, bp.keys_alt);
      std::swap(bp.vals, bp.vals_alt);
    } else {
      B2G_HIP(hipMemcpyAsync(keys, k0, n * sizeof(unsigned), hipMemcpyDeviceToDevice, w->stream));
      B2G_HIP(hipMemcpyAsync(vals, v0, n * sizeof(unsigned), hipMemcpyDeviceToDevice, w->stream));
    }
  }
  B2G_HIP(hipGetLastError());
  return B2G_OK;
}

}  // namespace b2g
