// Internal declarations shared by the FlashKNN HIP sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

#include "flashknn.h"

namespace fk {

constexpr int kWave = 64;
// k above this runs one wave per query (fk_query) and is refused by the
// batched kernel
constexpr int kMaxLaneK = 64;
// K classes of the batched launch (k up to 8, 16, 32, 64), one persistent
// kernel and one work queue each
constexpr int kBatchedClasses = 4;

// Slots of the small device scalar block owned by every grid handle.
enum Scalar : int {
  kStatus = 0,        // first error raised by a kernel (fk_status), 0 = ok
  kTotalCells = 1,    // cells of all per-cloud boxes
  kMaxDepth = 2,      // column layouts: largest box depth nz of a cloud
  kQueueHead = 3,     // batched mode: next work item of K class c at kQueueHead + c
  kPendBase = 7,      // batched mode: first pending-list slot of K class c at kPendBase + c
  kCounters = 16,     // query counter sets: kCounters + kCntStride * set
  kNumScalars = 48
};

// One set of query counters.  Plain queries alternate between sets 0 and 1
// (each first pass clears the other set for the next call, so no separate
// reset launch is needed); batched launches use sets 2 .. 5, one per K
// class, cleared by k_items.  (The certified count is not kept: it is the
// query count minus the pending one.)
enum Counter : int {
  kCntPending = 0,    // exact mode: queries left uncertified by pass 1
  kCntInline = 1,     // exact mode: uncertified ones the first pass expanded itself
  kCntExpanded = 2,   // exact mode: resolved by shell expansion (both kinds)
  kCntStride = 4      // slots per set (slot 3 spare)
};
constexpr int kBatchedSet = 2;  // K class c of a batched launch: set kBatchedSet + c
static_assert(kPendBase + kBatchedClasses <= kCounters, "scalar slots");
static_assert(kCounters + kCntStride * (kBatchedSet + kBatchedClasses) <= kNumScalars,
              "counter sets");
// Uncertified queries a plain first pass may expand itself per launch
// (fk_query.hip, run_tile); every claim counts in kCntInline, those past the
// budget are listed for the expansion launch instead.
#ifndef FK_INLINE_BUDGET
#define FK_INLINE_BUDGET 8
#endif
constexpr int kInlineBudget = FK_INLINE_BUDGET;

// Cell box of one cloud.  Dense layout: cell (lx,ly,lz) (local coordinates,
// 0 <= l < n) has table entry base + (lx*ny + ly)*nz + lz, so the three cells
// of a column segment (x, y, z-1..z+1) are adjacent and their points form one
// contiguous run of the cell-sorted arrays.  Column layout (boxes too large
// for one entry per cell): column (lx, ly) has entry base + lx*ny + ly, the
// points of a column are sorted by lz, and a segment's run is found by
// binary search over the sorted local z of the column.
struct CloudInfo {
  int32_t base;            // first table entry of the box
  int32_t nx, ny, nz;      // box size in cells (0 => empty cloud)
  int32_t minx, miny, minz;// absolute cell coordinate of local (0,0,0)
  int32_t first_id;        // offsets-built grids: first point id of the cloud
  float inv_cell;          // xyz-built grids: 1 / cell size (0 otherwise)
  int32_t num_points;      // points of the cloud
  int32_t columns;         // 1: column layout
  int32_t pad1;
};
static_assert(sizeof(CloudInfo) == 48, "CloudInfo layout");

// Per-cloud bounding boxes (6 ints: lo xyz, hi xyz) are accumulated with
// atomics.  With few clouds every block of a large cloud would update the
// same cache line, and such atomics serialise; blocks spread over
// box_copies(B) copies (each on its own 128-byte line) and readers take the
// union of the copies (the adaptive path's 9-cloud, 829k-point build: 165 ->
// 73 us).
__host__ __device__ __forceinline__ int box_copies(int64_t B) { return B <= 1024 ? 16 : 1; }
__host__ __device__ __forceinline__ int64_t box_stride(int64_t B) {
  return (6 * B + 31) / 32 * 32;
}
// Words the boxes of B clouds need, non-decreasing in B: a buffer sized for
// B clouds (a reserve) also holds the 16 copies of any smaller count.
__host__ __device__ __forceinline__ int64_t box_words(int64_t B) {
  const int64_t small = 16 * box_stride(1024);
  if (B <= 1024) return box_copies(B) * box_stride(B);
  return box_stride(B) > small ? box_stride(B) : small;
}

// Table entries of a cloud's box.
__host__ __device__ __forceinline__ int32_t table_entries(const CloudInfo &ci) {
  return ci.columns ? ci.nx * ci.ny : ci.nx * ci.ny * ci.nz;
}

// Table entry of local cell (lx, ly, lz) inside the box.
__device__ __forceinline__ int32_t lin_of(const CloudInfo &ci, int lx, int ly,
                                          int lz) {
  return ci.columns ? ci.base + lx * ci.ny + ly
                    : ci.base + (lx * ci.ny + ly) * ci.nz + lz;
}

__device__ __forceinline__ void raise_status(int32_t *scalars, int code) {
  atomicCAS(scalars + kStatus, 0, code);
}

// Rank within a counter (counts[v]++ for every active lane, returning the
// old value) aggregated over a whole block of kRankBlock threads: an LDS
// hash table groups the block's lanes by value, each distinct value takes
// one global atomic for its lanes, and a lane's rank is that base plus its
// order in the LDS count (ranks within a value are unique, in no fixed
// order).  Atomics on one address from different CUs serialise; coarse grids
// put thousands of points into one cell (one atomic per block and cell
// instead of one per wave took the adaptive path's 829k-point coarse-level
// count from 73 to 20 us).  Every thread of the block calls it; values are
// >= 0; lanes with act = false get 0.
constexpr int kRankBlock = 1024;
constexpr int kRankSlots = 2 * kRankBlock;  // > distinct values: probing ends

__device__ __forceinline__ int32_t block_rank(int32_t *counts, int32_t v,
                                              bool act) {
  __shared__ int32_t key[kRankSlots], cnt[kRankSlots], base[kRankSlots];
  for (int j = threadIdx.x; j < kRankSlots; j += blockDim.x) {
    key[j] = -1;
    cnt[j] = 0;
  }
  __syncthreads();
  int slot = 0, local = 0;
  if (act) {
    slot = (int)(((uint32_t)v * 2654435761u) >> 21);  // 11 bits
    for (;;) {
      const int old = atomicCAS(&key[slot], -1, v);
      if (old == -1 || old == v) break;
      slot = (slot + 1) & (kRankSlots - 1);
    }
    local = atomicAdd(&cnt[slot], 1);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kRankSlots; j += blockDim.x)
    if (key[j] >= 0) base[j] = atomicAdd(counts + key[j], cnt[j]);
  __syncthreads();
  return act ? base[slot] + local : 0;
}

__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  // Workgroups are dispatched round-robin over the 8 XCDs (bid % 8).  Give
  // each XCD one contiguous range of logical tiles so that neighbouring
  // tiles, which share candidate cells, share that XCD's L2.
  const int x = bid & 7, l = bid >> 3, per = nb >> 3, rem = nb & 7;
  return x * per + (x < rem ? x : rem) + l;
}

// A host-side device query (occupancy, CU count) cached per device: such
// calls cost microseconds of host time, and a query issues a few per call.
// Tag separates the cached quantities; f(dev, &value) computes one.
constexpr int kCacheDevices = 64;
template <class Tag, class F>
inline hipError_t cached_per_device(F &&f, int *out) {
  static std::atomic<int> cache[kCacheDevices];  // value + 1; 0 = not yet known
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const bool slot = dev >= 0 && dev < kCacheDevices;
  if (slot) {
    const int v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) {
      *out = v - 1;
      return hipSuccess;
    }
  }
  int r = 0;
  e = f(dev, &r);
  if (e != hipSuccess) return e;
  if (slot) cache[dev].store(r + 1, std::memory_order_relaxed);
  *out = r;
  return hipSuccess;
}
struct CuCountTag {};
inline hipError_t device_cus(int *cus) {
  return cached_per_device<CuCountTag>(
      [](int dev, int *r) {
        return hipDeviceGetAttribute(r, hipDeviceAttributeMultiprocessorCount, dev);
      },
      cus);
}

// ------------------------------------------------ Morton keys (16 bits/axis)
// Bit i of v moves to bit 3i (the standard magic-number interleave), and
// back.  A key holds x at bit 3i+2, y at 3i+1, z at 3i and the batch above
// bit 48 (the reference's xyz2key layout, FlashKNN/functions/z_order.py).
__host__ __device__ __forceinline__ uint64_t spread3(uint64_t v) {
  v &= 0xffffull;
  v = (v | (v << 16)) & 0x0000ff0000ffull;
  v = (v | (v << 8)) & 0x00f00f00f00full;
  v = (v | (v << 4)) & 0x0c30c30c30c3ull;
  v = (v | (v << 2)) & 0x249249249249ull;
  return v;
}
__host__ __device__ __forceinline__ uint32_t compact3(uint64_t v) {
  v &= 0x249249249249ull;
  v = (v | (v >> 2)) & 0x0c30c30c30c3ull;
  v = (v | (v >> 4)) & 0x00f00f00f00full;
  v = (v | (v >> 8)) & 0x0000ff0000ffull;
  v = (v | (v >> 16)) & 0xffffull;
  return (uint32_t)v;
}
__host__ __device__ __forceinline__ uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
}

// ------------------------------------------- adaptive neighbourhoods
// (fk_adaptive.hip builds, the query kernel is in fk_query.hip)
//
// Points sorted by key = batch << 48 | Morton(shifted coordinates): the
// cell of a point at level s (coordinates >> s) is key >> 3s, and the points
// of one cell of any level are one run of the sorted order.  Each level
// keeps an open-addressing table of 16-byte entries, cell key -> (run start,
// run length), ad_slots(n) entries per level (load <= 1/2), plus the cell's
// 27-cell candidate count in a parallel array.
constexpr int kAdLevels = 17;        // shifts 0..16
constexpr int kAdMaxBatch = 32768;   // batch indices in [0, 2^15)
constexpr uint64_t kAdEmpty = ~0ull;
constexpr int kAdStats = 4 + 4 * kAdLevels;
// stats (int64, device): status, reason, largest shifted coordinate, the
// coarsest level used, then per level s: points that selected s, and the
// sum / min / max of their 27-cell counts
enum AdStat : int { kAdStatus = 0, kAdReason = 1, kAdGmax = 2, kAdSmax = 3, kAdLevel0 = 4 };
constexpr int kAdReasonBatch = 1;    // a batch index outside [0, 2^15)
constexpr int kAdReasonSpan = 2;     // a batch spans 2^16 cells or more on an axis

struct alignas(16) AdEntry {
  uint64_t key;    // cell key at the entry's level, kAdEmpty when free
  int32_t start;   // first sorted position of the cell's run
  int32_t count;   // run length
};

__host__ __device__ __forceinline__ int ad_log2_slots(int64_t n) {
  int lg = 6;
  while ((int64_t(1) << lg) < 2 * n) ++lg;
  return lg;
}
__host__ __device__ __forceinline__ int64_t ad_slots(int64_t n) {
  return int64_t(1) << ad_log2_slots(n);
}
// Home slot of cell key c in a table of 2^lg entries: Fibonacci hashing,
// the top bits of c * 2^64 / phi.  (Keeping the 8 cells of a 2x2x2 block in
// one 128-byte line made absent neighbours probe through full lines: the
// 27-cell counts of a 92k-point sweep took 214 instead of 127 us.)
__host__ __device__ __forceinline__ int64_t ad_home(uint64_t c, int lg) {
  return (int64_t)(((c ^ (c >> 31)) * 0x9e3779b97f4a7c15ull) >> (64 - lg));
}
// Entry index of cell c at level s, or -1 (the table always has a free slot).
__device__ __forceinline__ int64_t ad_find(const AdEntry *__restrict__ ent, int lg, int s,
                                           uint64_t c) {
  const int64_t P = int64_t(1) << lg;
  const AdEntry *t = ent + (int64_t)s * P;
  int64_t h = ad_home(c, lg);
  for (;;) {
    const uint64_t k = t[h].key;
    if (k == c) return (int64_t)s * P + h;
    if (k == kAdEmpty) return -1;
    h = (h + 1) & (P - 1);
  }
}
// N cells of level s at once: the first probes are independent 16-byte
// loads, all in flight together (the rare longer chains follow); entries of
// absent cells (and of c = kAdEmpty) come back as {kAdEmpty, 0, 0}.
template <int N>
__device__ __forceinline__ void ad_find_n(const AdEntry *__restrict__ ent, int lg, int s,
                                          const uint64_t (&c)[N], AdEntry (&out)[N],
                                          int64_t (&slot)[N]) {
  const int64_t P = int64_t(1) << lg;
  const AdEntry *t = ent + (int64_t)s * P;
  int64_t h[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    h[j] = ad_home(c[j], lg);
    out[j] = c[j] == kAdEmpty ? AdEntry{kAdEmpty, 0, 0} : t[h[j]];
  }
#pragma unroll
  for (int j = 0; j < N; ++j) {
    while (out[j].key != c[j] && out[j].key != kAdEmpty) {
      h[j] = (h[j] + 1) & (P - 1);
      out[j] = t[h[j]];
    }
    const bool hit = c[j] != kAdEmpty && out[j].key == c[j];
    slot[j] = hit ? (int64_t)s * P + h[j] : -1;
    if (!hit) out[j] = AdEntry{kAdEmpty, 0, 0};
  }
}

// Cells of level s around cell c: decode once, spread each axis's three
// coordinates once, and combine (27 keys from 9 spreads); kAdEmpty where a
// coordinate leaves [0, 2^(16-s)).  Neighbour j is (dx, dy, dz) =
// (j / 9 - 1, j / 3 % 3 - 1, j % 3 - 1).
struct AdNeighbours {
  uint64_t hi;            // batch bits of the level's key
  uint64_t sp[3][3];      // spread coordinates per axis and offset
  bool ok[3][3];
  __device__ __forceinline__ void init(uint64_t c, int s) {
    const int mb = 48 - 3 * s;
    const uint64_t m = mb > 0 ? (c & ((1ull << mb) - 1)) : 0ull;
    hi = c - m;
    const int lim = 1 << (16 - s);
    const int v[3] = {(int)compact3(m >> 2), (int)compact3(m >> 1), (int)compact3(m)};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int x = v[a] + d - 1;
        ok[a][d] = x >= 0 && x < lim;
        sp[a][d] = spread3((uint64_t)(ok[a][d] ? x : 0)) << (2 - a);
      }
  }
  __device__ __forceinline__ uint64_t key(int j) const {
    const int dx = j / 9, dy = (j / 3) % 3, dz = j % 3;
    return (ok[0][dx] && ok[1][dy] && ok[2][dz]) ? (hi | sp[0][dx] | sp[1][dy] | sp[2][dz])
                                                 : kAdEmpty;
  }
};

struct AdaptiveArgs {
  const uint64_t *skeys;   // [n] sorted keys
  const float4 *srec;      // [n] sorted (x, y, z, input index bits)
  const int32_t *level;    // [n] selected level of each sorted point
  const AdEntry *ent;      // [kAdLevels * P]
  const int32_t *hc27;     // [kAdLevels * P]
  const int64_t *stats;    // [kAdStats]
  int32_t lg;              // P = 2^lg entries per level
  int32_t n;
  int32_t k;
  float cut_r2;
  int64_t *out_idx;        // [n, k] rows in input order
  float *out_d;            // [n, k] squared distances (nullable)
};
// The 27-cell query at each point's selected level, K class KC (fk_query.hip).
template <int KC>
hipError_t launch_adaptive_query(const AdaptiveArgs &a, hipStream_t s);
// Sort, tables, counts and level selection (fk_adaptive.hip); the tables of
// a workspace it filled, for the query.
int64_t adaptive_workspace_bytes(int64_t n);
hipError_t launch_adaptive_select(const int64_t *grid, const int64_t *batch,
                                  const float *xyz, int64_t n, int32_t lo, int32_t hi,
                                  int32_t max_shift, void *ws, int32_t *out_level,
                                  int64_t *stats, hipStream_t s);
hipError_t adaptive_tables(void *ws, int64_t n, const int64_t *stats, AdaptiveArgs *args);

// --------------------------------------------------------------- launches
// (defined in fk_build.hip / fk_query.hip / fk_backward.hip)

hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s);
hipError_t launch_morton_encode(const int64_t *grid, const int64_t *batch,
                                int64_t n, int64_t *keys, hipStream_t s);
hipError_t launch_morton_neighbors(const int64_t *grid_down,
                                   const int64_t *pbatch, int64_t P,
                                   int64_t *keys, hipStream_t s);

struct BuildBuffers {
  int32_t *scalars;      // [kNumScalars]
  int32_t *bbox;         // [6 * clouds]
  CloudInfo *clouds;     // [clouds]
  int4 *pt_cell;         // [n]   absolute cell + cloud of input points
  int32_t *pt_lin;       // [n]   linear cell id of input points
  int32_t *pt_rank;      // [n]   rank inside the cell (atomic order)
  int32_t *counts;       // [cap_cells]
  int32_t *cell_start;   // [cap_cells + 1]
  int32_t *scan_part;    // [scan blocks]
  float4 *srec;          // [n]   sorted (x, y, z, id bits)
  int4 *scell;           // [n]   sorted local cell + cloud
  int64_t cap_cells;
  // column layout only: radix sort of (table entry, lz) keys
  uint64_t *ckeys[2];    // [n] each
  int32_t *cperm[2];     // [n] each
  void *sort_tmp;
  size_t sort_tmp_bytes;
};

hipError_t launch_build_cells(const BuildBuffers &b, const float *xyz,
                              const int64_t *grid, const int64_t *batch,
                              int64_t n, int32_t num_clouds, int32_t num_down,
                              hipStream_t s);
hipError_t launch_build_xyz_cells(const BuildBuffers &b, const float *xyz,
                                  const int32_t *offsets, int32_t num_clouds,
                                  int64_t n, const float *inv_cell,
                                  hipStream_t s);
hipError_t launch_auto_cell(int32_t *bbox, const float *xyz,
                            const int32_t *offsets, int32_t B, int64_t n,
                            const int32_t *k_per_cloud, int32_t k, float scale,
                            float *side_out, float *inv_out, hipStream_t s);
// columns: 0 dense boxes, 1 column layout (see CloudInfo)
hipError_t launch_layout(const BuildBuffers &b, int32_t num_clouds,
                         const int32_t *offsets, const float *inv_cell,
                         int32_t columns, hipStream_t s);
// count / scan / scatter; cells_hint = host upper bound of total cells
hipError_t launch_sort(const BuildBuffers &b, const float *xyz,
                       const int64_t *grid, int64_t n, int64_t cells_hint,
                       hipStream_t s);
// Column layout: counts and start table as launch_sort, then the points
// ordered by (table entry, lz) with a radix sort of key_bits-bit keys
// (entry << depth_bits | lz).  (fk_columns.hip)
hipError_t launch_sort_columns(const BuildBuffers &b, const float *xyz,
                               const int64_t *grid, int64_t n,
                               int64_t cells_hint, int depth_bits, int key_bits,
                               hipStream_t s);
// Temporary bytes of the radix sort of n pairs with key_bits-bit keys.
hipError_t column_sort_tmp_bytes(int64_t n, int key_bits, size_t *bytes);
// Counting-sort steps shared by both layouts (fk_build.hip): counts per
// table entry, in-entry ranks, start table.
hipError_t launch_count_scan(const BuildBuffers &b, int64_t n,
                             int64_t cells_hint, hipStream_t s);

struct QueryBuffers {
  float4 *qrec;          // [nq] sorted (x, y, z, output row bits)
  int4 *qcell;           // [nq] sorted local cell + cloud
  int32_t *qpad;         // [nq] sorted per-query pad value
  int32_t *q_lin;        // [nq]
  int32_t *q_rank;       // [nq]
  int4 *q_cell_abs;      // [nq]
  int32_t *qcounts;      // [cap_cells]
  int32_t *qstart;       // [cap_cells + 1]
};

hipError_t launch_attach(const BuildBuffers &b, const QueryBuffers &q,
                         const float *qxyz, const int64_t *qcell,
                         const int64_t *qbatch, const int32_t *qoffsets,
                         int64_t nq, const int32_t *qpad, int32_t num_clouds,
                         int32_t num_down, int from_xyz, int64_t cells_hint,
                         hipStream_t s);

struct QueryArgs {
  const float4 *srec;
  const int4 *scell;     // sorted local cells (column layouts search their z)
  const int32_t *cell_start;
  const CloudInfo *clouds;
  const float4 *qrec;
  const int4 *qcell;
  const int32_t *qpad;  // nullable
  int32_t nq;
  int32_t k;
  int32_t R;
  float cut_r2;
  int32_t pad_mode;
  int32_t dist_mode;
  float pad_dist;
  int32_t idx64;
  void *out_idx;
  float *out_d;
  int32_t exact;
  int32_t passes;        // bit 0 reset, bit 1 first pass, bit 2 expansion
  int32_t prune;         // FK_QUERY_PRUNE: prune neighbour cells mid-scan
  int32_t one_lane;      // FK_QUERY_ONE_LANE: one lane per query (64 per wave)
  int32_t batch_sel;     // candidate batch: 0 default, 1 FK_QUERY_BATCH_SMALL,
                         // 2 FK_QUERY_BATCH_LARGE
  int32_t columns;       // the grid uses the column layout (see CloudInfo)
  int32_t *scalars;
  int32_t *qcnt;         // this call's counter set (Counter slots)
  int32_t *qcnt_next;    // set the first pass clears for the next call (or null)
  int32_t *pend_q;       // [nq]
  uint64_t *pend_keys;   // [nq * k]
  // batched mode: the K class whose pending list (from slot
  // scalars[kPendBase + pend_class] on) this launch appends to or expands;
  // -1: the list starts at slot 0
  int32_t pend_class;
  // batched mode
  const int32_t *k_per_cloud;
  const int64_t *out_offsets;
  int32_t num_clouds;
  int32_t num_items;     // work items (cloud, 64-query tile)
  const int32_t *item_cloud;   // unused when 0 items
  // balanced first pass (self queries): per-wave (first query, lanes per
  // query) of a launch of plan_waves waves; null: uniform 32-query tiles
  const int2 *plan;
  int32_t plan_waves;
};

hipError_t launch_query(const QueryArgs &a, hipStream_t s);

// Balanced first pass.  The queries are cut into super-tiles of
// kPlanTile consecutive (cell-sorted) queries; a super-tile whose largest
// block candidate count is c runs on G = clamp(pow2ceil(c / L), 1, 8) waves
// (G lanes per query, kWave / G queries per wave), with L the smallest
// per-lane budget for which the waves fit `capacity`, heaviest super-tiles
// first.  A uniform launch is set by its slowest tiles (dense regions hold
// several times the median candidate count and each lane scans its records
// one round after another); the plan gives those more lanes and light ones
// fewer.  Results do not depend on the plan.
constexpr int kPlanTile = 64;

// Resident waves of the planned first pass of K class kc and candidate
// batch batch_sel (QueryArgs::batch_sel) on this device.
hipError_t planned_capacity(int kc, int batch_sel, int32_t *waves);
// Writes plan[0, capacity) (first query, lanes); tile_cost: int32[ceil(nq /
// kPlanTile)] scratch.  Needs nq <= kPlanTile * capacity.
hipError_t launch_plan(const QueryArgs &a, int32_t capacity, int32_t *tile_cost,
                       int2 *plan, hipStream_t s);
hipError_t launch_plan_columns(const QueryArgs &a, int32_t capacity,
                               int32_t *tile_cost, int2 *plan, hipStream_t s);
// First pass of K class KC (8, 16, 32, 64) on the dense / column layout;
// each is instantiated in its class's translation unit of fk_query.hip.
template <int KC>
hipError_t launch_first_pass(const QueryArgs &a, hipStream_t s);
template <int KC>
hipError_t launch_first_pass_columns(const QueryArgs &a, hipStream_t s);
// Persistent batched first pass over the work items of K class KC (queue
// and prefix slot cls); launch_batched_first dispatches on cls.
template <int KC>
hipError_t launch_batched_class(const QueryArgs &a, int cls, hipStream_t s);
template <int KC>
hipError_t launch_batched_class_columns(const QueryArgs &a, int cls, hipStream_t s);
hipError_t launch_batched_first(const QueryArgs &a, int cls, hipStream_t s);
hipError_t launch_batched_first_columns(const QueryArgs &a, int cls, hipStream_t s);
// Resident blocks per CU of the planned first pass of K class KC.
template <int KC>
hipError_t planned_blocks(int batch_sel, int *nb);
// Shell expansion of the pending (uncertified) queries of an exact launch,
// and the first pass of k > kMaxLaneK (one wave per query), per layout.
hipError_t launch_expand(const QueryArgs &a, hipStream_t s);
hipError_t launch_expand_columns(const QueryArgs &a, hipStream_t s);
hipError_t launch_wide(const QueryArgs &a, hipStream_t s);
hipError_t launch_wide_columns(const QueryArgs &a, hipStream_t s);

int kclass_of(int k);  // 8, 16, 32, 64 or 128
// The query launches built for the column layout (fk_query.hip compiled with
// FK_COLUMNS=1); the launches above forward to them when a.columns is set.
hipError_t launch_query_columns(const QueryArgs &a, hipStream_t s);
hipError_t launch_candidate_count_columns(const QueryArgs &a,
                                          unsigned long long *total,
                                          hipStream_t s);
hipError_t launch_point_candidates_columns(const QueryArgs &a, int32_t *out,
                                           hipStream_t s);
hipError_t launch_cloud_candidates_columns(const QueryArgs &a,
                                           unsigned long long *acc,
                                           hipStream_t s);
hipError_t launch_adapt_cell_columns(const QueryArgs &a, unsigned long long *acc,
                                     const int32_t *k_per_cloud, int32_t k,
                                     float target_ratio,
                                     const float *ratio_per_cloud,
                                     float exponent, float *side_io,
                                     float *inv_io, hipStream_t s);
hipError_t launch_candidate_count(const QueryArgs &a, unsigned long long *total,
                                  hipStream_t s);
hipError_t launch_point_candidates(const QueryArgs &a, int32_t *out,
                                   hipStream_t s);
// Per-cloud candidate sums of the self queries (per_cloud: u64[num_clouds],
// zeroed here), then side *= clamp((target_ratio*k / mean)^(1/exponent))
// and inv = 1/side, in place on the device.
hipError_t launch_adapt_cell(const QueryArgs &a, unsigned long long *per_cloud,
                             const int32_t *k_per_cloud, int32_t k,
                             float target_ratio, const float *ratio_per_cloud,
                             float exponent, float *side_io, float *inv_io,
                             hipStream_t s);
hipError_t launch_cloud_candidates(const QueryArgs &a, unsigned long long *acc,
                                   hipStream_t s);
// Per-class work-item prefixes of a batched launch ([kBatchedClasses][B +
// 1]: an item is one tile of a cloud of that K class), queue heads reset;
// a k_per_cloud entry outside [1, max_k] raises FK_ERR_K_RANGE.
hipError_t launch_items(const CloudInfo *clouds, const int32_t *cell_start,
                        const int32_t *k_per_cloud, int32_t B, int32_t max_k,
                        int32_t *prefix, int32_t *scalars, int32_t *qcnt,
                        hipStream_t s);
int64_t scan_parts_needed(int64_t n);
hipError_t launch_scan_host(const int32_t *in, int32_t *out, int64_t n,
                            int32_t *part, hipStream_t s);

// backward
// voxel-first subsampling
int64_t voxel_workspace_bytes(int64_t n);
hipError_t launch_voxel_first(const float *xyz, const int32_t *offsets,
                              int32_t B, int64_t n, float size, void *ws,
                              int64_t *out_idx, int64_t *out_count,
                              hipStream_t s);
// The same over the library's static per-device workspace (n at most
// FK_VOXEL_NOALLOC_MAX_POINTS; hipErrorInvalidValue above).  Allocates nothing.
hipError_t launch_voxel_first_noalloc(const float *xyz, const int32_t *offsets,
                                      int32_t B, int64_t n, float size,
                                      int64_t *out_idx, int64_t *out_count,
                                      hipStream_t s);

int64_t backward_workspace_bytes(int64_t nq, int64_t ns, int32_t k);
hipError_t launch_dist_backward(const float *q, int64_t nq, const float *p,
                                int64_t ns, const int64_t *idx, int32_t k,
                                const float *grad_d, float *grad_q,
                                float *grad_p, void *ws, hipStream_t s);
hipError_t launch_idw_forward(const float *feat, int64_t ns, int32_t C,
                              const int64_t *idx, const float *w, int64_t nq,
                              int32_t k, float *out, hipStream_t s);
hipError_t launch_idw_backward(const float *feat, int64_t ns, int32_t C,
                               const int64_t *idx, const float *w, int64_t nq,
                               int32_t k, const float *grad_out,
                               float *grad_feat, float *grad_w, void *ws,
                               hipStream_t s);
// fused IDW interpolation (distances, normalised weights, gather; backward
// to features, queries and supports over one reverse CSR)
int64_t interp_workspace_bytes(int64_t nq, int64_t ns, int32_t k);
hipError_t launch_interp_forward(const float *q, int64_t nq, const float *p,
                                 int64_t ns, const int64_t *idx, int32_t k,
                                 const float *feat, int32_t C, float eps,
                                 float *w_out, float *out, hipStream_t s);
hipError_t launch_interp_backward(const float *q, int64_t nq, const float *p,
                                  int64_t ns, const int64_t *idx, int32_t k,
                                  const float *feat, int32_t C, float eps,
                                  const float *grad_out, float *grad_feat,
                                  float *grad_q, float *grad_p, void *ws,
                                  hipStream_t s);

}  // namespace fk
