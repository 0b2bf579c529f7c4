// Adaptive neighbourhoods on the device: per-point octree level selection
// (AdaptiveNeighborhoodFlashKNN.query, FlashKNN/functions/adaptive_octree.py
// :491-612; normalisation :129-146, node statistic :204-208, selection rule
// :312-351).
//
// The reference builds a Morton octree with neighbour propagation and
// flattens the selected levels into one query launch.  Here one radix sort
// of key = batch << 48 | Morton(shifted coordinates) serves every level: the
// cell of a point at level s (coordinates >> s) is key >> 3s, so the points
// of any cell of any level form one run of the sorted order.  Each level
// keeps an open-addressing table, cell key -> (run start, run length,
// 27-cell candidate count); a run start inserts its cell, the 27-cell count
// of a cell is the sum of its neighbours' run lengths, and the coarse-to-
// fine walk of a point reads its own cell's count per level.  Nothing is
// read back to the host; invalid batch indices or spans raise a status in
// the stats block.  The query of the selected levels is in fk_query.hip.
#include "fk_internal.h"

#include <rocprim/device/device_radix_sort.hpp>

namespace fk {

namespace {

constexpr int kBlock = 256;
constexpr int64_t kSortBits = 63;  // 15 batch bits + 48 Morton bits

inline unsigned blocks_of(int64_t n) {
  const int64_t b = (n + kBlock - 1) / kBlock;
  return (unsigned)(b < 1 ? 1 : b);
}

// Workspace carve (byte offsets, 256-aligned).
struct AdLayout {
  int64_t mins, keys0, keys1, vals0, vals1, srec, level, ent, hc27, tmp;
  int64_t total;
  size_t tmp_bytes;
};

inline int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

hipError_t ad_layout(int64_t n, AdLayout &L) {
  const int64_t m = n < 1 ? 1 : n;
  const int64_t LP = (int64_t)kAdLevels * ad_slots(n);
  size_t tmp = 0;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, tmp, (const uint64_t *)nullptr,
                                           (uint64_t *)nullptr, (const int32_t *)nullptr,
                                           (int32_t *)nullptr, (unsigned)m, 0u,
                                           (unsigned)kSortBits, (hipStream_t)0);
  if (e != hipSuccess) return e;
  int64_t o = 0;
  auto take = [&](int64_t bytes) {
    const int64_t at = o;
    o = align256(o + bytes);
    return at;
  };
  L.mins = take(8 * 3 * (int64_t)kAdMaxBatch);
  L.keys0 = take(8 * m);
  L.keys1 = take(8 * m);
  L.vals0 = take(4 * m);
  L.vals1 = take(4 * m);
  L.srec = take(16 * m);
  L.level = take(4 * m);
  L.ent = take((int64_t)sizeof(AdEntry) * LP);
  L.hc27 = take(4 * LP);
  L.tmp = take((int64_t)tmp);
  L.tmp_bytes = tmp;
  L.total = o;
  return hipSuccess;
}

__device__ __forceinline__ void ad_raise(int64_t *stats, int reason) {
  atomicCAS((unsigned long long *)(stats + kAdStatus), 0ull,
            (unsigned long long)FK_ERR_DOMAIN);
  atomicCAS((unsigned long long *)(stats + kAdReason), 0ull, (unsigned long long)reason);
}

// Per-batch minima (as order-preserving unsigned: x ^ 2^63), the table
// entries (empty), the stats block.
__global__ void k_ad_init(uint64_t *mins, AdEntry *ent, int64_t LP, int64_t *stats) {
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = t0; i < 3 * (int64_t)kAdMaxBatch; i += stride) mins[i] = ~0ull;
  for (int64_t i = t0; i < LP; i += stride) ent[i] = AdEntry{kAdEmpty, 0, 0};
  if (t0 < kAdStats) {
    const int r = (int)t0 - kAdLevel0;
    stats[t0] = (r >= 0 && (r & 3) == 2) ? INT64_MAX : 0;  // per-level minima
  }
}

__device__ __forceinline__ uint64_t bias(int64_t v) { return (uint64_t)v ^ (1ull << 63); }
__device__ __forceinline__ int64_t unbias(uint64_t v) { return (int64_t)(v ^ (1ull << 63)); }

// Block-wide reductions (kBigBlock threads, every thread calls).
constexpr int kBigBlock = 1024;
constexpr int kMinPer = 1;  // points per thread of the minimum kernel

template <class T, class Op>
__device__ __forceinline__ T block_reduce(T v, Op op, T *lds) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, kWave));
  __syncthreads();  // lds free from an earlier reduction
  if (lane == 0) lds[w] = v;
  __syncthreads();
  v = lds[0];
  for (int j = 1; j < (int)(blockDim.x >> 6); ++j) v = op(v, lds[j]);
  return v;
}

// Minimum of each axis per batch.  A block whose points all lie in one
// batch (points of a batch are usually stored together) takes one atomic per
// axis; other blocks one per point.  (One per wave on one batch serialised
// 4,300 atomics on three addresses: 51 us for 92k points.)
__global__ void __launch_bounds__(kBigBlock)
k_ad_min(const int64_t *__restrict__ grid, const int64_t *__restrict__ batch, int64_t n,
         uint64_t *mins, int64_t *stats) {
  __shared__ int64_t red_i[16];
  __shared__ uint64_t red_u[16];
  const int64_t base = (int64_t)blockIdx.x * kBigBlock * kMinPer + threadIdx.x;
  int64_t bmin = INT64_MAX, bmax = INT64_MIN;
  uint64_t v[3] = {~0ull, ~0ull, ~0ull};
  for (int r = 0; r < kMinPer; ++r) {
    const int64_t i = base + (int64_t)r * kBigBlock;
    if (i >= n) break;
    const int64_t b = batch ? batch[i] : 0;
    if (b < 0 || b >= kAdMaxBatch) {
      ad_raise(stats, kAdReasonBatch);
      continue;
    }
    bmin = b < bmin ? b : bmin;
    bmax = b > bmax ? b : bmax;
    for (int a = 0; a < 3; ++a) {
      const uint64_t g = bias(grid[3 * i + a]);
      v[a] = g < v[a] ? g : v[a];
    }
  }
  auto mn_i = [](int64_t x, int64_t y) { return x < y ? x : y; };
  auto mx_i = [](int64_t x, int64_t y) { return x > y ? x : y; };
  auto mn_u = [](uint64_t x, uint64_t y) { return x < y ? x : y; };
  bmin = block_reduce(bmin, mn_i, red_i);
  bmax = block_reduce(bmax, mx_i, red_i);
  if (bmin > bmax) return;  // no valid point in the block (uniform)
  if (bmin == bmax) {
    for (int a = 0; a < 3; ++a) {
      const uint64_t m = block_reduce(v[a], mn_u, red_u);
      if (threadIdx.x == 0)
        atomicMin((unsigned long long *)&mins[3 * bmin + a], (unsigned long long)m);
    }
    return;
  }
  for (int r = 0; r < kMinPer; ++r) {
    const int64_t i = base + (int64_t)r * kBigBlock;
    if (i >= n) break;
    const int64_t b = batch ? batch[i] : 0;
    if (b < 0 || b >= kAdMaxBatch) continue;
    for (int a = 0; a < 3; ++a)
      atomicMin((unsigned long long *)&mins[3 * b + a], (unsigned long long)bias(grid[3 * i + a]));
  }
}

// key = batch << 48 | Morton(coordinates - batch minimum); the largest
// shifted coordinate (the default coarsest level is its bit length), one
// atomic per block.
__global__ void __launch_bounds__(kBigBlock)
k_ad_keys(const int64_t *__restrict__ grid, const int64_t *__restrict__ batch, int64_t n,
          const uint64_t *__restrict__ mins, uint64_t *keys, int32_t *vals, int64_t *stats) {
  __shared__ int64_t red[16];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t gm = 0;
  if (i < n) {
    const int64_t b = batch ? batch[i] : 0;
    uint64_t key = kAdEmpty;
    if (b >= 0 && b < kAdMaxBatch) {
      int64_t g[3];
      bool span_ok = true;
      for (int a = 0; a < 3; ++a) {
        g[a] = grid[3 * i + a] - unbias(mins[3 * b + a]);
        span_ok &= g[a] >= 0 && g[a] < 65536;
        gm = g[a] > gm ? g[a] : gm;
      }
      if (span_ok)
        key = ((uint64_t)b << 48) | morton3((uint32_t)g[0], (uint32_t)g[1], (uint32_t)g[2]);
      else
        ad_raise(stats, kAdReasonSpan);
    }
    keys[i] = key;
    vals[i] = (int32_t)i;
  }
  gm = block_reduce(gm, [](int64_t x, int64_t y) { return x > y ? x : y; }, red);
  if (threadIdx.x == 0 && gm > 0)
    atomicMax((unsigned long long *)(stats + kAdGmax), (unsigned long long)gm);
}

__device__ __forceinline__ int ad_smax(const int64_t *stats, int max_shift) {
  if (max_shift >= 0) return max_shift;
  const int64_t g = stats[kAdGmax];
  return g > 0 ? 64 - __clzll((unsigned long long)g) : 0;
}

// Sorted position p: the point's record (its coordinates, or the shifted
// cell coordinates as floats when no xyz is given, and its input index).
__global__ void k_ad_records(const float *__restrict__ xyz, const int64_t *__restrict__ grid,
                             const int64_t *__restrict__ batch, int64_t n,
                             const uint64_t *__restrict__ mins,
                             const int32_t *__restrict__ perm, float4 *srec,
                             int64_t *stats, int max_shift) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p == 0) stats[kAdSmax] = ad_smax(stats, max_shift);
  if (p >= n || stats[kAdStatus] != 0) return;
  const int64_t i = perm[p];
  float x, y, z;
  if (xyz) {
    x = xyz[3 * i];
    y = xyz[3 * i + 1];
    z = xyz[3 * i + 2];
  } else {
    const int64_t b = batch ? batch[i] : 0;
    x = (float)(grid[3 * i] - unbias(mins[3 * b]));
    y = (float)(grid[3 * i + 1] - unbias(mins[3 * b + 1]));
    z = (float)(grid[3 * i + 2] - unbias(mins[3 * b + 2]));
  }
  srec[p] = make_float4(x, y, z, __int_as_float((int)i));
}

// Point and level of a thread of a (points, levels) launch: blockIdx.y is
// the level; along x each XCD takes one contiguous range of points of every
// level (the cells whose table entries its L2 holds are then neighbours, and
// the XCDs share every level's work alike).
__device__ __forceinline__ bool ad_item(int64_t n, int64_t &p, int &s) {
  s = (int)blockIdx.y;
  p = xcd_remap(blockIdx.x, gridDim.x) * (int64_t)blockDim.x + threadIdx.x;
  return p < n;
}

// A point that starts its cell's run at level s (the previous point lies in
// another cell) inserts the cell with the run's first position.
__global__ void k_ad_insert(const uint64_t *__restrict__ keys, int64_t n, int lg,
                            AdEntry *ent, const int64_t *stats) {
  int64_t p;
  int s;
  if (!ad_item(n, p, s) || stats[kAdStatus] != 0 || s > (int)stats[kAdSmax]) return;
  const uint64_t c = keys[p] >> (3 * s);
  if (p > 0 && (keys[p - 1] >> (3 * s)) == c) return;
  const int64_t P = int64_t(1) << lg;
  AdEntry *t = ent + (int64_t)s * P;
  int64_t h = ad_home(c, lg);
  while (atomicCAS((unsigned long long *)&t[h].key, (unsigned long long)kAdEmpty,
                   (unsigned long long)c) != (unsigned long long)kAdEmpty)
    h = (h + 1) & (P - 1);
  t[h].start = (int32_t)p;
}

// The last point of a run writes the run's length.
__global__ void k_ad_ends(const uint64_t *__restrict__ keys, int64_t n, int lg, AdEntry *ent,
                          const int64_t *stats) {
  int64_t p;
  int s;
  if (!ad_item(n, p, s) || stats[kAdStatus] != 0 || s > (int)stats[kAdSmax]) return;
  const uint64_t c = keys[p] >> (3 * s);
  if (p + 1 < n && (keys[p + 1] >> (3 * s)) == c) return;
  const int64_t slot = ad_find(ent, lg, s, c);
  ent[slot].count = (int32_t)(p + 1 - ent[slot].start);
}

// 27-cell candidate count of every cell of every level: the run lengths of
// the cell and its 26 neighbours within the batch (one thread per run start
// and level, the 27 lookups in flight together).
__global__ void k_ad_count27(const uint64_t *__restrict__ keys, int64_t n, int lg,
                             const AdEntry *__restrict__ ent, int32_t *hc27,
                             const int64_t *stats) {
  int64_t p;
  int s;
  if (!ad_item(n, p, s) || stats[kAdStatus] != 0 || s > (int)stats[kAdSmax]) return;
  const uint64_t c = keys[p] >> (3 * s);
  if (p > 0 && (keys[p - 1] >> (3 * s)) == c) return;
  AdNeighbours nb;
  nb.init(c, s);
  uint64_t nc[27];
#pragma unroll
  for (int j = 0; j < 27; ++j) nc[j] = nb.key(j);
  AdEntry e[27];
  int64_t slot[27];
  ad_find_n<27>(ent, lg, s, nc, e, slot);
  int total = 0;
#pragma unroll
  for (int j = 0; j < 27; ++j) total += e[j].count;
  hc27[slot[13]] = total;  // j = 13: the cell itself
}

// The coarse-to-fine walk of every point: start at the coarsest level; keep
// descending while the count exceeds hi, and take a finer level only while
// its count is at least lo.  The point's cells of all levels are looked up
// at once.  Per-level statistics of the selected counts are aggregated per
// block in LDS first.
__global__ void __launch_bounds__(kBlock)
k_ad_select(const uint64_t *__restrict__ keys, const float4 *__restrict__ srec, int64_t n,
            int lg, const AdEntry *__restrict__ ent, const int32_t *__restrict__ hc27,
            int32_t lo, int32_t hi, int32_t *level, int32_t *out_level, int64_t *stats) {
  __shared__ unsigned long long s_cnt[kAdLevels], s_sum[kAdLevels], s_min[kAdLevels],
      s_max[kAdLevels];
  if (threadIdx.x < kAdLevels) {
    s_cnt[threadIdx.x] = 0;
    s_sum[threadIdx.x] = 0;
    s_min[threadIdx.x] = ~0ull;
    s_max[threadIdx.x] = 0;
  }
  __syncthreads();
  const int64_t p = xcd_remap(blockIdx.x, gridDim.x) * (int64_t)blockDim.x + threadIdx.x;
  if (p < n && stats[kAdStatus] == 0) {
    const int smax = (int)stats[kAdSmax];
    const uint64_t key = keys[p];
    // own cell of every level: first probes all in flight, then the chains
    const int64_t P = int64_t(1) << lg;
    uint64_t cs[kAdLevels], k[kAdLevels];
    int64_t h[kAdLevels];
#pragma unroll
    for (int s = 0; s < kAdLevels; ++s) {
      cs[s] = s <= smax ? key >> (3 * s) : kAdEmpty;
      h[s] = ad_home(cs[s], lg);
      k[s] = s <= smax ? ent[(int64_t)s * P + h[s]].key : kAdEmpty;
    }
    int c27[kAdLevels];
#pragma unroll
    for (int s = 0; s < kAdLevels; ++s) {
      while (k[s] != cs[s] && k[s] != kAdEmpty) {
        h[s] = (h[s] + 1) & (P - 1);
        k[s] = ent[(int64_t)s * P + h[s]].key;
      }
      c27[s] = s <= smax ? hc27[(int64_t)s * P + h[s]] : 0;
    }
    int sel = smax;
    int csel = 0;
#pragma unroll
    for (int s = 0; s < kAdLevels; ++s)
      if (s == smax) csel = c27[s];
    bool active = csel > hi;
    for (int s = smax - 1; s >= 0 && active; --s) {
      int c = 0;
#pragma unroll
      for (int j = 0; j < kAdLevels; ++j)
        if (j == s) c = c27[j];
      if (c < lo) break;
      sel = s;
      csel = c;
      active = c > hi;
    }
    level[p] = sel;
    if (out_level) out_level[__float_as_int(srec[p].w)] = sel;
    atomicAdd(&s_cnt[sel], 1ull);
    atomicAdd(&s_sum[sel], (unsigned long long)csel);
    atomicMin(&s_min[sel], (unsigned long long)csel);
    atomicMax(&s_max[sel], (unsigned long long)csel);
  }
  __syncthreads();
  if (threadIdx.x < kAdLevels && s_cnt[threadIdx.x]) {
    int64_t *st = stats + kAdLevel0 + 4 * threadIdx.x;
    atomicAdd((unsigned long long *)st, s_cnt[threadIdx.x]);
    atomicAdd((unsigned long long *)(st + 1), s_sum[threadIdx.x]);
    atomicMin((unsigned long long *)(st + 2), s_min[threadIdx.x]);
    atomicMax((unsigned long long *)(st + 3), s_max[threadIdx.x]);
  }
}

}  // namespace

int64_t adaptive_workspace_bytes(int64_t n) {
  AdLayout L;
  if (ad_layout(n, L) != hipSuccess) return -1;
  return L.total;
}

hipError_t adaptive_tables(void *ws, int64_t n, const int64_t *stats, AdaptiveArgs *args) {
  AdLayout L;
  hipError_t e = ad_layout(n, L);
  if (e != hipSuccess) return e;
  char *w = (char *)ws;
  args->skeys = (const uint64_t *)(w + L.keys1);
  args->srec = (const float4 *)(w + L.srec);
  args->level = (const int32_t *)(w + L.level);
  args->ent = (const AdEntry *)(w + L.ent);
  args->hc27 = (const int32_t *)(w + L.hc27);
  args->stats = stats;
  args->lg = ad_log2_slots(n);
  args->n = (int32_t)n;
  return hipSuccess;
}

hipError_t launch_adaptive_select(const int64_t *grid, const int64_t *batch,
                                  const float *xyz, int64_t n, int32_t lo, int32_t hi,
                                  int32_t max_shift, void *ws, int32_t *out_level,
                                  int64_t *stats, hipStream_t s) {
  AdLayout L;
  hipError_t e = ad_layout(n, L);
  if (e != hipSuccess) return e;
  char *w = (char *)ws;
  uint64_t *mins = (uint64_t *)(w + L.mins);
  uint64_t *keys0 = (uint64_t *)(w + L.keys0), *keys1 = (uint64_t *)(w + L.keys1);
  int32_t *vals0 = (int32_t *)(w + L.vals0), *vals1 = (int32_t *)(w + L.vals1);
  float4 *srec = (float4 *)(w + L.srec);
  int32_t *level = (int32_t *)(w + L.level);
  AdEntry *ent = (AdEntry *)(w + L.ent);
  int32_t *hc27 = (int32_t *)(w + L.hc27);
  const int lg = ad_log2_slots(n);
  const int64_t LP = (int64_t)kAdLevels * (int64_t(1) << lg);
  hipLaunchKernelGGL(k_ad_init, dim3(blocks_of(LP > 3 * kAdMaxBatch ? LP : 3 * kAdMaxBatch)),
                     dim3(kBlock), 0, s, mins, ent, LP, stats);
  const int64_t per_min = (int64_t)kBigBlock * kMinPer;
  hipLaunchKernelGGL(k_ad_min, dim3((unsigned)((n + per_min - 1) / per_min)), dim3(kBigBlock),
                     0, s, grid, batch, n, mins, stats);
  hipLaunchKernelGGL(k_ad_keys, dim3((unsigned)((n + kBigBlock - 1) / kBigBlock)),
                     dim3(kBigBlock), 0, s, grid, batch, n, mins, keys0, vals0, stats);
  size_t tmp = L.tmp_bytes;
  e = rocprim::radix_sort_pairs(w + L.tmp, tmp, keys0, keys1, vals0, vals1, (unsigned)n, 0u,
                                (unsigned)kSortBits, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_ad_records, dim3(blocks_of(n)), dim3(kBlock), 0, s, xyz, grid, batch,
                     n, mins, vals1, srec, stats, max_shift);
  // one thread per (point, level)
  const dim3 nl(blocks_of(n), kAdLevels);
  hipLaunchKernelGGL(k_ad_insert, nl, dim3(kBlock), 0, s, keys1, n, lg, ent, stats);
  hipLaunchKernelGGL(k_ad_ends, nl, dim3(kBlock), 0, s, keys1, n, lg, ent, stats);
  hipLaunchKernelGGL(k_ad_count27, nl, dim3(kBlock), 0, s, keys1, n, lg, ent, hc27, stats);
  hipLaunchKernelGGL(k_ad_select, dim3(blocks_of(n)), dim3(kBlock), 0, s, keys1, srec, n, lg,
                     ent, hc27, lo, hi, level, out_level, stats);
  return hipGetLastError();
}

}  // namespace fk
