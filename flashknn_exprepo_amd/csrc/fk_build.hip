// Grid construction for gfx950: cell ids, per-cloud dense cell boxes,
// counting sort of points by cell, cell start table, sorted records.
//
// Replaces the torch construction of FlashKNN.query
// (FlashKNN/functions/FlashKnnWrapper.py:226-337: Morton key, argsort,
// unique_consecutive, searchsorted neighbour table, cumsum candidate table).
// Instead of a Morton sort plus a searched 27-neighbour table, every cloud
// gets a dense box of cells laid out x-major, z-fastest: the 27-cell
// neighbourhood of a cell is 9 column segments, each one contiguous run of
// the cell-sorted points, found by two reads of the cell start table.  No
// host synchronisation is needed when the cell capacity is reserved.
#include "fk_internal.h"

#include <climits>

namespace fk {

namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;  // 4096 elements per block

__device__ __forceinline__ bool wave_all(bool p) {
  return __ballot(!p) == 0ull;  // inactive lanes count as true
}

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Fill p[0, *n_dev + extra) with v (size read on the device).
__global__ void k_fill_dev(int32_t *p, int32_t v, const int32_t *n_dev,
                           int32_t extra) {
  const int64_t n = (int64_t)(*n_dev) + extra;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ------------------------------------------------------------------ Morton

// (spread3: fk_internal.h)
__device__ __forceinline__ int64_t morton(int64_t x, int64_t y, int64_t z,
                                          int64_t b) {
  return (int64_t)((spread3((uint64_t)x) << 2) | (spread3((uint64_t)y) << 1) |
                   spread3((uint64_t)z) | ((uint64_t)b << 48));
}

__global__ void k_morton_encode(const int64_t *grid, const int64_t *batch,
                                int64_t n, int64_t *keys) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = morton(grid[3 * i], grid[3 * i + 1], grid[3 * i + 2],
                   batch ? batch[i] : 0);
}

__global__ void k_morton_neighbors(const int64_t *gd, const int64_t *pb,
                                   int64_t P, int64_t *keys) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= P * 27) return;
  const int64_t p = t / 27;
  const int j = (int)(t - p * 27);
  // offsets [0,-1,1] per axis, x slowest (ij meshgrid order, centre first)
  const int ox = j / 9, oy = (j / 3) % 3, oz = j % 3;
  auto off = [](int s) { return s == 0 ? 0 : (s == 1 ? -1 : 1); };
  keys[t] = morton(gd[3 * p] + off(ox), gd[3 * p + 1] + off(oy),
                   gd[3 * p + 2] + off(oz), pb ? pb[p] : 0);
}

// ---------------------------------------------------------- cells + bbox

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
  return v;
}

// Per-cloud bounding boxes.  Each thread accumulates kItems points, the wave
// and then the block reduce, so a block whose points all belong to one cloud
// issues six atomics in total (same-address atomics serialise in L2: one
// set per wave cost ~0.1 ms per 100k points).  Values are ints; float
// coordinates go through the monotone mapping ord_of.  One point per thread:
// the cell kernel over 829k points (9 clouds) takes 21 us with one, 37 us
// with two and 70 us with four (the box copies, 16 to 256, change nothing).
constexpr int kItems = 1;
constexpr int kMaxWaves = 16;

struct Box {
  int cloud;  // -1: nothing accumulated yet
  int lo[3], hi[3];
};

__device__ __forceinline__ void box_init(Box &b) {
  b.cloud = -1;
  b.lo[0] = b.lo[1] = b.lo[2] = INT_MAX;
  b.hi[0] = b.hi[1] = b.hi[2] = INT_MIN;
}

__device__ __forceinline__ void box_atomic(int32_t *bbox, int c, const int *lo,
                                           const int *hi) {
  int32_t *p = bbox + 6 * c;
  atomicMin(p + 0, lo[0]); atomicMin(p + 1, lo[1]); atomicMin(p + 2, lo[2]);
  atomicMax(p + 3, hi[0]); atomicMax(p + 4, hi[1]); atomicMax(p + 5, hi[2]);
}

// A point of a cloud other than the one being accumulated goes straight to
// memory (happens only at cloud boundaries).
__device__ __forceinline__ void box_add(Box &b, int32_t *bbox, int c, int x,
                                        int y, int z) {
  if (b.cloud < 0) b.cloud = c;
  if (c == b.cloud) {
    b.lo[0] = min(b.lo[0], x); b.lo[1] = min(b.lo[1], y); b.lo[2] = min(b.lo[2], z);
    b.hi[0] = max(b.hi[0], x); b.hi[1] = max(b.hi[1], y); b.hi[2] = max(b.hi[2], z);
  } else {
    const int v[3] = {x, y, z};
    box_atomic(bbox, c, v, v);
  }
}

// Called by every thread of the block.
__device__ void box_flush_block(Box &b, int32_t *bbox) {
  __shared__ int s_cloud[kMaxWaves];
  __shared__ int s_lo[kMaxWaves][3], s_hi[kMaxWaves][3];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const bool has = b.cloud >= 0;
  const uint64_t m = __ballot(has);
  int wc = -1;
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  if (m != 0ull) {
    const int c0 = __shfl(b.cloud, (int)__builtin_ctzll(m), kWave);
    if (wave_all(!has || b.cloud == c0)) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(has ? b.lo[a] : INT_MAX);
        hi[a] = wave_max(has ? b.hi[a] : INT_MIN);
      }
      wc = c0;
    } else if (has) {
      box_atomic(bbox, b.cloud, b.lo, b.hi);
    }
  }
  if (lane == 0) {
    s_cloud[w] = wc;
#pragma unroll
    for (int a = 0; a < 3; ++a) { s_lo[w][a] = lo[a]; s_hi[w][a] = hi[a]; }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int cur = -1;
    int clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
    for (int j = 0; j < nw; ++j) {
      const int c = s_cloud[j];
      if (c < 0) continue;
      if (c != cur) {
        if (cur >= 0) box_atomic(bbox, cur, clo, chi);
        cur = c;
#pragma unroll
        for (int a = 0; a < 3; ++a) { clo[a] = s_lo[j][a]; chi[a] = s_hi[j][a]; }
      } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          clo[a] = min(clo[a], s_lo[j][a]);
          chi[a] = max(chi[a], s_hi[j][a]);
        }
      }
    }
    if (cur >= 0) box_atomic(bbox, cur, clo, chi);
  }
}

// Monotone float -> int map (for non-NaN values) and its inverse.
__device__ __forceinline__ int ord_of(float f) {
  const int i = __float_as_int(f);
  return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ __forceinline__ float float_of_ord(int i) {
  return __int_as_float(i >= 0 ? i : (i ^ 0x7fffffff));
}

// The copy of the per-cloud boxes a block accumulates into.
__device__ __forceinline__ int32_t *box_copy(int32_t *bbox, int32_t B) {
  return bbox + (int64_t)(blockIdx.x & (box_copies(B) - 1)) * box_stride(B);
}

// Union of the copies of cloud c's box -> b[6].
__device__ __forceinline__ void box_read(const int32_t *bbox, int32_t B, int c,
                                         int32_t *b) {
  const int64_t stride = box_stride(B);
  const int32_t *p = bbox + 6 * c;
  b[0] = p[0]; b[1] = p[1]; b[2] = p[2]; b[3] = p[3]; b[4] = p[4]; b[5] = p[5];
  for (int k = 1; k < box_copies(B); ++k) {
    const int32_t *q = p + k * stride;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      b[a] = min(b[a], q[a]);
      b[3 + a] = max(b[3 + a], q[3 + a]);
    }
  }
}

__global__ void k_init_bbox(int32_t *bbox, int32_t B) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= (int64_t)box_copies(B) * B) return;
  int32_t *p = bbox + (e / B) * box_stride(B) + 6 * (e % B);
  p[0] = p[1] = p[2] = INT_MAX;
  p[3] = p[4] = p[5] = INT_MIN;
}

constexpr int64_t kCellLimit = 1ll << 30;  // |cell coordinate| bound

__device__ __forceinline__ bool cell_ok(int64_t v) {
  return v > -kCellLimit && v < kCellLimit;
}

// cell = grid >> num_down (arithmetic shift = floor division by 2^num_down)
__global__ void __launch_bounds__(256)
k_cells_from_grid(const int64_t *grid, const int64_t *batch, int64_t n,
                  int32_t num_down, int32_t B, int4 *pt_cell, int32_t *bbox_all,
                  int32_t *scalars) {
  int32_t *bbox = box_copy(bbox_all, B);
  Box box;
  box_init(box);
  const int64_t tile = (int64_t)blockIdx.x * blockDim.x * kItems;
  // all loads first: the atomics below would otherwise keep the compiler
  // from overlapping one item's loads with the previous item's work
  int64_t c[kItems][3], bt[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t i = tile + threadIdx.x + (int64_t)j * blockDim.x;
    c[j][0] = c[j][1] = c[j][2] = bt[j] = 0;
    if (i < n) {
      c[j][0] = grid[3 * i]; c[j][1] = grid[3 * i + 1]; c[j][2] = grid[3 * i + 2];
      bt[j] = batch ? batch[i] : 0;
    }
  }
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t i = tile + threadIdx.x + (int64_t)j * blockDim.x;
    if (i < n) {
      const int64_t gx = c[j][0] >> num_down, gy = c[j][1] >> num_down,
                    gz = c[j][2] >> num_down;
      const int64_t b = bt[j];
      if (b < 0 || b >= B || !cell_ok(gx) || !cell_ok(gy) || !cell_ok(gz)) {
        raise_status(scalars, FK_ERR_DOMAIN);
        pt_cell[i] = make_int4(0, 0, 0, -1);
      } else {
        pt_cell[i] = make_int4((int)gx, (int)gy, (int)gz, (int)b);
        box_add(box, bbox, (int)b, (int)gx, (int)gy, (int)gz);
      }
    }
  }
  box_flush_block(box, bbox);
}

// first c with offsets[c] > i  (offsets = cumulative segment ends)
__device__ __forceinline__ int cloud_of(const int32_t *offsets, int B,
                                        int64_t i) {
  int lo = 0, hi = B;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((int64_t)offsets[mid] > i) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// Scaled coordinate u = x * inv (one rounding, no contraction possible);
// the query kernels recompute it identically for certification.
__device__ __forceinline__ float scaled(float x, float inv) {
  return __fmul_rn(x, inv);
}

__global__ void __launch_bounds__(256)
k_cells_from_xyz(const float *xyz, const int32_t *offsets, int32_t B, int64_t n,
                 const float *inv_cell, int4 *pt_cell, int32_t *bbox_all,
                 int32_t *scalars) {
  int32_t *bbox = box_copy(bbox_all, B);
  Box box;
  box_init(box);
  const int64_t tile = (int64_t)blockIdx.x * blockDim.x * kItems;
  // loads first (see k_cells_from_grid)
  float p[kItems][3];
  int cls[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t i = tile + threadIdx.x + (int64_t)j * blockDim.x;
    p[j][0] = p[j][1] = p[j][2] = 0.f;
    cls[j] = B;
    if (i < n) {
      p[j][0] = xyz[3 * i]; p[j][1] = xyz[3 * i + 1]; p[j][2] = xyz[3 * i + 2];
      cls[j] = cloud_of(offsets, B, i);
    }
  }
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t i = tile + threadIdx.x + (int64_t)j * blockDim.x;
    if (i < n) {
      const int cl = cls[j];
      bool ok = cl < B;
      float fx = 0.f, fy = 0.f, fz = 0.f;
      if (ok) {
        const float inv = inv_cell[cl];
        fx = floorf(scaled(p[j][0], inv));
        fy = floorf(scaled(p[j][1], inv));
        fz = floorf(scaled(p[j][2], inv));
        // rejects NaN as well
        ok = fabsf(fx) < (float)kCellLimit && fabsf(fy) < (float)kCellLimit &&
             fabsf(fz) < (float)kCellLimit;
      }
      if (!ok) {
        raise_status(scalars, FK_ERR_DOMAIN);
        pt_cell[i] = make_int4(0, 0, 0, -1);
      } else {
        pt_cell[i] = make_int4((int)fx, (int)fy, (int)fz, cl);
        box_add(box, bbox, cl, (int)fx, (int)fy, (int)fz);
      }
    }
  }
  box_flush_block(box, bbox);
}

// Float bounding box of every cloud (ord_of-mapped ints in bbox[6c..6c+5]).
__global__ void __launch_bounds__(256)
k_cloud_bounds(const float *xyz, const int32_t *offsets, int32_t B, int64_t n,
               int32_t *bbox_all) {
  int32_t *bbox = box_copy(bbox_all, B);
  Box box;
  box_init(box);
  const int64_t tile = (int64_t)blockIdx.x * blockDim.x * kItems;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t i = tile + threadIdx.x + (int64_t)j * blockDim.x;
    if (i < n) {
      const int cl = cloud_of(offsets, B, i);
      const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
      if (cl < B && x == x && y == y && z == z)
        box_add(box, bbox, cl, ord_of(x), ord_of(y), ord_of(z));
    }
  }
  box_flush_block(box, bbox);
}

// Cell side per cloud from its bounding-box density:
// side = scale * (k * volume / n)^(1/3), degenerate axes counted with 1% of
// the largest extent, and at least (volume / (64 n))^(1/3) so that the dense
// cell box holds at most ~64 cells per point.
__global__ void k_auto_cell(const int32_t *bbox, const int32_t *offsets,
                            int32_t B, const int32_t *k_per_cloud, int32_t k,
                            float scale, float *side_out, float *inv_out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= B) return;
  int32_t b[6];
  box_read(bbox, B, c, b);
  const int64_t n = offsets[c] - (c ? offsets[c - 1] : 0);
  double side = 1.0;
  if (n > 0 && b[0] <= b[3]) {
    double ext[3], span = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      ext[a] = (double)float_of_ord(b[3 + a]) - (double)float_of_ord(b[a]);
      span = fmax(span, ext[a]);
    }
    span = fmax(span, 1e-6);
    double vol = 1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) vol *= fmax(ext[a], 1e-2 * span);
    const double kk = (double)(k_per_cloud ? k_per_cloud[c] : k);
    side = (double)scale * cbrt(kk * vol / (double)n);
    side = fmax(side, cbrt(vol / (64.0 * (double)n)));
    side = fmax(side, 1e-6);
  }
  const float s = (float)side;
  if (side_out) side_out[c] = s;
  inv_out[c] = (float)(1.0 / (double)s);
}

// One block: per-cloud box sizes, exclusive scan of table entries (cells,
// or columns in the column layout), CloudInfo.
__global__ void __launch_bounds__(1024)
k_layout(const int32_t *bbox, int32_t B, const int32_t *offsets,
         const float *inv_cell, int64_t cap_cells, int32_t columns,
         CloudInfo *clouds, int32_t *scalars) {
  __shared__ int64_t warp_sums[16];
  __shared__ int64_t carry;
  __shared__ int max_depth;
  if (threadIdx.x == 0) {
    carry = 0;
    max_depth = 0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int c0 = 0; c0 < B; c0 += 1024) {
    const int c = c0 + threadIdx.x;
    int64_t cells = 0;
    int nx = 0, ny = 0, nz = 0, mnx = 0, mny = 0, mnz = 0;
    if (c < B) {
      int32_t b[6];
      box_read(bbox, B, c, b);
      if (b[0] <= b[3]) {
        mnx = b[0]; mny = b[1]; mnz = b[2];
        const int64_t ex = (int64_t)b[3] - b[0] + 1, ey = (int64_t)b[4] - b[1] + 1,
                      ez = (int64_t)b[5] - b[2] + 1;
        // ez < 2^31 (|cell| < 2^30), so nz fits an int in both layouts
        const int64_t entries = columns ? ex * ey : ex * ey * ez;
        if (entries > kCellLimit) {
          raise_status(scalars, FK_ERR_GRID_TOO_LARGE);
        } else {
          nx = (int)ex; ny = (int)ey; nz = (int)ez;
          cells = entries;
          if (columns) atomicMax(&max_depth, nz);
        }
      }
    }
    // inclusive wave scan
    int64_t v = cells;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int64_t u = __shfl_up(v, o, kWave);
      if (lane >= o) v += u;
    }
    if (lane == 63) warp_sums[w] = v;
    __syncthreads();
    int64_t before = carry;
    for (int j = 0; j < w; ++j) before += warp_sums[j];
    const int64_t excl = before + v - cells;
    if (c < B) {
      CloudInfo ci;
      ci.base = (int32_t)(excl < INT_MAX ? excl : INT_MAX);
      ci.nx = nx; ci.ny = ny; ci.nz = nz;
      ci.minx = mnx; ci.miny = mny; ci.minz = mnz;
      ci.first_id = offsets ? (c ? offsets[c - 1] : 0) : 0;
      ci.num_points = offsets ? offsets[c] - ci.first_id : 0;
      ci.inv_cell = inv_cell ? inv_cell[c] : 0.f;
      ci.columns = columns;
      ci.pad1 = 0;
      clouds[c] = ci;
    }
    __syncthreads();
    if (threadIdx.x == 1023) carry = excl + cells;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int64_t total = carry;
    scalars[kMaxDepth] = max_depth;
    if (total > cap_cells) {
      raise_status(scalars, FK_ERR_GRID_TOO_LARGE);
      scalars[kTotalCells] = (int32_t)(cap_cells < INT_MAX ? cap_cells : INT_MAX);
    } else {
      scalars[kTotalCells] = (int32_t)total;
    }
  }
}

// In-cell rank from a counting atomic, aggregated over the lanes of a wave
// that fall in one cell: the first kAggRounds distinct cells of the wave
// take one atomic each (coarse grids put whole waves in one cell, whose
// per-lane atomics would serialise), the remaining lanes one atomic each.
// Every lane of the wave calls it; lanes with act = false get 0.
constexpr int kAggRounds = 4;

__device__ __forceinline__ int32_t wave_rank(int32_t *counts, int32_t lin,
                                             bool act) {
  const int lane = threadIdx.x & 63;
  uint64_t pending = __ballot(act);
  int32_t rank = 0;
  for (int r = 0; r < kAggRounds && pending; ++r) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int32_t llin = __shfl(lin, leader, kWave);
    const uint64_t same = __ballot(act && lin == llin) & pending;
    int32_t base = 0;
    if (lane == leader) base = atomicAdd(counts + llin, (int)__popcll(same));
    base = __shfl(base, leader, kWave);
    if ((same >> lane) & 1)
      rank = base + (int)__popcll(same & ((1ull << lane) - 1));
    pending &= ~same;
  }
  if ((pending >> lane) & 1) rank = atomicAdd(counts + lin, 1);
  return rank;
}

// Table entry and in-entry rank of every point.
__global__ void __launch_bounds__(kRankBlock)
k_count(const int4 *pt_cell, int64_t n, const CloudInfo *clouds,
        int32_t *counts, int32_t *pt_lin, int32_t *pt_rank,
        const int32_t *scalars) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool act = i < n && scalars[kStatus] == 0;
  int32_t lin = 0;
  if (act) {
    const int4 c = pt_cell[i];
    const CloudInfo ci = clouds[c.w];
    lin = lin_of(ci, c.x - ci.minx, c.y - ci.miny, c.z - ci.minz);
    pt_lin[i] = lin;
  }
  const int32_t rank = block_rank(counts, lin, act);
  if (act) pt_rank[i] = rank;
}

// ------------------------------------------------------------------- scan
// Exclusive scan out[0..N] of in[0..N), N = *n_dev + extra, out[N] = total.

__device__ __forceinline__ int block_reduce_sum(int v, int *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  int s = 0;
  for (int j = 0; j < (int)(blockDim.x >> 6); ++j) s += red[j];
  __syncthreads();
  return s;
}

__global__ void __launch_bounds__(kBlock)
k_scan_reduce(const int32_t *in, const int32_t *n_dev, int64_t n_fixed,
              int32_t extra, int32_t *part) {
  __shared__ int red[kBlock / 64];
  const int64_t N = (n_dev ? (int64_t)(*n_dev) : n_fixed) + extra;
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int64_t idx = base + threadIdx.x + j * kBlock;
    if (idx < N) s += in[idx];
  }
  s = block_reduce_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(1024)
k_scan_partials(int32_t *part, int32_t nparts) {
  __shared__ int wsum[16];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int c0 = 0; c0 < nparts; c0 += 1024) {
    const int c = c0 + threadIdx.x;
    const int x = c < nparts ? part[c] : 0;
    int v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(v, o, kWave);
      if (lane >= o) v += u;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    int before = carry;
    for (int j = 0; j < w; ++j) before += wsum[j];
    if (c < nparts) part[c] = before + v - x;
    __syncthreads();
    if (threadIdx.x == 1023) carry = before + v;
    __syncthreads();
  }
}

// LDS index of tile element i, padded by one word per 16 so that both the
// striped (i = j * kBlock + t) and the blocked (i = t * kScanItems + j)
// accesses of a wave are free of bank conflicts.
constexpr __host__ __device__ int scan_pad(int i) { return i + (i >> 4); }

// The tile goes through LDS: loaded and stored striped (consecutive lanes,
// consecutive words: coalesced), scanned blocked (kScanItems consecutive
// elements per thread).  Reading and writing each thread's own 16
// consecutive words straight from memory made every wave access 64
// separate 64-byte strides (~1.8 TB/s on a 38M-entry table).
__global__ void __launch_bounds__(kBlock)
k_scan_apply(const int32_t *in, const int32_t *n_dev, int64_t n_fixed,
             int32_t extra, const int32_t *part, int32_t *out) {
  __shared__ int tile[scan_pad(kScanTile)];
  __shared__ int wsum[kBlock / 64];
  const int64_t N = (n_dev ? (int64_t)(*n_dev) : n_fixed) + extra;
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int i = j * kBlock + t;
    const int64_t idx = base + i;
    tile[scan_pad(i)] = idx < N ? in[idx] : 0;
  }
  __syncthreads();
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = tile[scan_pad(t * kScanItems + j)];
    s += v[j];
  }
  const int lane = t & 63, w = t >> 6;
  int incl = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int before = part[blockIdx.x];
  for (int j = 0; j < w; ++j) before += wsum[j];
  int run = before + incl - s;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    tile[scan_pad(t * kScanItems + j)] = run;
    run += v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int i = j * kBlock + t;
    const int64_t idx = base + i;
    if (idx <= N) out[idx] = tile[scan_pad(i)];
  }
}

// ---------------------------------------------------------------- scatter

__global__ void k_scatter(const float *xyz, const int64_t *grid, int64_t n,
                          const int4 *pt_cell, const int32_t *pt_lin,
                          const int32_t *pt_rank, const int32_t *cell_start,
                          const CloudInfo *clouds, float4 *srec, int4 *scell,
                          const int32_t *scalars) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n || scalars[kStatus] != 0) return;
  const int4 c = pt_cell[i];
  const CloudInfo ci = clouds[c.w];
  const int32_t pos = cell_start[pt_lin[i]] + pt_rank[i];
  float x, y, z;
  if (xyz) {
    x = xyz[3 * i]; y = xyz[3 * i + 1]; z = xyz[3 * i + 2];
  } else {
    x = (float)grid[3 * i]; y = (float)grid[3 * i + 1]; z = (float)grid[3 * i + 2];
  }
  srec[pos] = make_float4(x, y, z, __int_as_float((int)i));
  scell[pos] = make_int4(c.x - ci.minx, c.y - ci.miny, c.z - ci.minz, c.w);
}

// ---------------------------------------------------------------- queries

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

__device__ __forceinline__ int sat_local(int64_t v) {
  return (int)(v < -kCellLimit ? -kCellLimit : (v > kCellLimit ? kCellLimit : v));
}

// Query cell in the grid's local cell space, its sort bucket (the cell,
// clamped into the box; `total` for clouds without cells) and rank.
__global__ void k_qcells(const float *qxyz, const int64_t *qcell,
                         const int64_t *qbatch, const int32_t *qoffsets,
                         int64_t nq, int32_t B, int32_t num_down,
                         int32_t from_xyz, const CloudInfo *clouds,
                         int32_t *qcounts, int32_t *q_lin, int32_t *q_rank,
                         int4 *q_cell, int32_t *scalars) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  bool act = i < nq && scalars[kStatus] == 0;
  int cl = 0;
  int64_t ax = 0, ay = 0, az = 0;
  if (act && from_xyz) {
    cl = cloud_of(qoffsets, B, i);
    if (cl >= B) {
      raise_status(scalars, FK_ERR_DOMAIN);
      act = false;
    } else {
      const float inv = clouds[cl].inv_cell;
      const float fx = floorf(scaled(qxyz[3 * i], inv));
      const float fy = floorf(scaled(qxyz[3 * i + 1], inv));
      const float fz = floorf(scaled(qxyz[3 * i + 2], inv));
      if (!(fabsf(fx) < (float)kCellLimit && fabsf(fy) < (float)kCellLimit &&
            fabsf(fz) < (float)kCellLimit)) {
        raise_status(scalars, FK_ERR_DOMAIN);
        act = false;
      }
      ax = (int64_t)fx; ay = (int64_t)fy; az = (int64_t)fz;
    }
  } else if (act) {
    const int64_t b = qbatch ? qbatch[i] : 0;
    if (b < 0 || b >= B) {
      raise_status(scalars, FK_ERR_DOMAIN);
      act = false;
    } else {
      cl = (int)b;
      ax = qcell[3 * i] >> num_down;
      ay = qcell[3 * i + 1] >> num_down;
      az = qcell[3 * i + 2] >> num_down;
    }
  }
  int32_t lin = 0;
  if (act) {
    const CloudInfo ci = clouds[cl];
    const int lx = sat_local(ax - ci.minx), ly = sat_local(ay - ci.miny),
              lz = sat_local(az - ci.minz);
    q_cell[i] = make_int4(lx, ly, lz, cl);
    if (ci.nx == 0)
      lin = scalars[kTotalCells];
    else
      lin = lin_of(ci, clampi(lx, 0, ci.nx - 1), clampi(ly, 0, ci.ny - 1),
                   clampi(lz, 0, ci.nz - 1));
    q_lin[i] = lin;
  }
  const int32_t rank = wave_rank(qcounts, lin, act);
  if (act) q_rank[i] = rank;
}

__global__ void k_qscatter(const float *qxyz, int64_t nq, const int32_t *q_lin,
                           const int32_t *q_rank, const int4 *q_cell,
                           const int32_t *qstart, const int32_t *qpad_in,
                           float4 *qrec, int4 *qcell, int32_t *qpad,
                           const int32_t *scalars) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nq || scalars[kStatus] != 0) return;
  const int32_t pos = qstart[q_lin[i]] + q_rank[i];
  qrec[pos] = make_float4(qxyz[3 * i], qxyz[3 * i + 1], qxyz[3 * i + 2],
                          __int_as_float((int)i));
  qcell[pos] = q_cell[i];
  qpad[pos] = qpad_in ? qpad_in[i] : -1;
}

inline unsigned blocks_for(int64_t n, int per = kBlock) {
  int64_t b = (n + per - 1) / per;
  return (unsigned)(b < 1 ? 1 : b);
}

hipError_t launch_scan(const int32_t *in, int32_t *out, const int32_t *n_dev,
                       int32_t extra, int64_t n_hint, int32_t *part,
                       hipStream_t s) {
  const unsigned nb = blocks_for(n_hint + extra + 1, kScanTile);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, s, in, n_dev,
                     n_hint, extra, part);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, part,
                     (int32_t)nb);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kBlock), 0, s, in, n_dev,
                     n_hint, extra, part, out);
  return hipGetLastError();
}

}  // namespace

int64_t scan_parts_needed(int64_t n) { return (n + 2 + kScanTile - 1) / kScanTile + 1; }

// Exclusive scan of a host-sized array: out[0..n] (out[n] = total).
hipError_t launch_scan_host(const int32_t *in, int32_t *out, int64_t n,
                            int32_t *part, hipStream_t s) {
  return launch_scan(in, out, nullptr, 0, n, part, s);
}

hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  unsigned nb = blocks_for(n);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(kBlock), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_morton_encode(const int64_t *grid, const int64_t *batch,
                                int64_t n, int64_t *keys, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton_encode, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                     grid, batch, n, keys);
  return hipGetLastError();
}

hipError_t launch_morton_neighbors(const int64_t *gd, const int64_t *pb,
                                   int64_t P, int64_t *keys, hipStream_t s) {
  if (P <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton_neighbors, dim3(blocks_for(P * 27)),
                     dim3(kBlock), 0, s, gd, pb, P, keys);
  return hipGetLastError();
}

hipError_t launch_build_cells(const BuildBuffers &b, const float *xyz,
                              const int64_t *grid, const int64_t *batch,
                              int64_t n, int32_t B, int32_t num_down,
                              hipStream_t s) {
  (void)xyz;
  hipLaunchKernelGGL(k_init_bbox, dim3(blocks_for((int64_t)box_copies(B) * B)), dim3(kBlock), 0, s,
                     b.bbox, B);
  if (n > 0)
    hipLaunchKernelGGL(k_cells_from_grid, dim3(blocks_for(n, kBlock * kItems)),
                       dim3(kBlock), 0, s, grid, batch, n, num_down, B,
                       b.pt_cell, b.bbox, b.scalars);
  return hipGetLastError();
}

hipError_t launch_build_xyz_cells(const BuildBuffers &b, const float *xyz,
                                  const int32_t *offsets, int32_t B, int64_t n,
                                  const float *inv_cell, hipStream_t s) {
  hipLaunchKernelGGL(k_init_bbox, dim3(blocks_for((int64_t)box_copies(B) * B)), dim3(kBlock), 0, s,
                     b.bbox, B);
  if (n > 0)
    hipLaunchKernelGGL(k_cells_from_xyz, dim3(blocks_for(n, kBlock * kItems)),
                       dim3(kBlock), 0, s, xyz, offsets, B, n, inv_cell,
                       b.pt_cell, b.bbox, b.scalars);
  return hipGetLastError();
}

hipError_t launch_auto_cell(int32_t *bbox, const float *xyz,
                            const int32_t *offsets, int32_t B, int64_t n,
                            const int32_t *k_per_cloud, int32_t k, float scale,
                            float *side_out, float *inv_out, hipStream_t s) {
  hipLaunchKernelGGL(k_init_bbox, dim3(blocks_for((int64_t)box_copies(B) * B)), dim3(kBlock), 0, s, bbox, B);
  if (n > 0)
    hipLaunchKernelGGL(k_cloud_bounds, dim3(blocks_for(n, kBlock * kItems)),
                       dim3(kBlock), 0, s, xyz, offsets, B, n, bbox);
  hipLaunchKernelGGL(k_auto_cell, dim3(blocks_for(B)), dim3(kBlock), 0, s, bbox,
                     offsets, B, k_per_cloud, k, scale, side_out, inv_out);
  return hipGetLastError();
}

hipError_t launch_layout(const BuildBuffers &b, int32_t B,
                         const int32_t *offsets, const float *inv_cell,
                         int32_t columns, hipStream_t s) {
  hipLaunchKernelGGL(k_layout, dim3(1), dim3(1024), 0, s, b.bbox, B, offsets,
                     inv_cell, b.cap_cells, columns, b.clouds, b.scalars);
  return hipGetLastError();
}

hipError_t launch_count_scan(const BuildBuffers &b, int64_t n,
                             int64_t cells_hint, hipStream_t s) {
  unsigned nfill = blocks_for(cells_hint + 1);
  if (nfill > 8192) nfill = 8192;
  hipLaunchKernelGGL(k_fill_dev, dim3(nfill), dim3(kBlock), 0, s, b.counts, 0,
                     b.scalars + kTotalCells, 0);
  if (n > 0)
    hipLaunchKernelGGL(k_count, dim3(blocks_for(n, kRankBlock)), dim3(kRankBlock), 0, s,
                       b.pt_cell, n, b.clouds, b.counts, b.pt_lin, b.pt_rank,
                       b.scalars);
  return launch_scan(b.counts, b.cell_start, b.scalars + kTotalCells, 0,
                     cells_hint, b.scan_part, s);
}

hipError_t launch_sort(const BuildBuffers &b, const float *xyz,
                       const int64_t *grid, int64_t n, int64_t cells_hint,
                       hipStream_t s) {
  hipError_t e = launch_count_scan(b, n, cells_hint, s);
  if (e != hipSuccess) return e;
  if (n > 0)
    hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n)), dim3(kBlock), 0, s, xyz,
                       grid, n, b.pt_cell, b.pt_lin, b.pt_rank, b.cell_start,
                       b.clouds, b.srec, b.scell, b.scalars);
  return hipGetLastError();
}

hipError_t launch_attach(const BuildBuffers &b, const QueryBuffers &q,
                         const float *qxyz, const int64_t *qcell,
                         const int64_t *qbatch, const int32_t *qoffsets,
                         int64_t nq, const int32_t *qpad, int32_t B,
                         int32_t num_down, int from_xyz, int64_t cells_hint,
                         hipStream_t s) {
  unsigned nfill = blocks_for(cells_hint + 2);
  if (nfill > 8192) nfill = 8192;
  // one extra bucket (index total) for queries of clouds without cells
  hipLaunchKernelGGL(k_fill_dev, dim3(nfill), dim3(kBlock), 0, s, q.qcounts, 0,
                     b.scalars + kTotalCells, 1);
  if (nq > 0)
    hipLaunchKernelGGL(k_qcells, dim3(blocks_for(nq)), dim3(kBlock), 0, s, qxyz,
                       qcell, qbatch, qoffsets, nq, B, num_down, from_xyz,
                       b.clouds, q.qcounts, q.q_lin, q.q_rank, q.q_cell_abs,
                       b.scalars);
  hipError_t e = launch_scan(q.qcounts, q.qstart, b.scalars + kTotalCells, 1,
                             cells_hint, b.scan_part, s);
  if (e != hipSuccess) return e;
  if (nq > 0)
    hipLaunchKernelGGL(k_qscatter, dim3(blocks_for(nq)), dim3(kBlock), 0, s,
                       qxyz, nq, q.q_lin, q.q_rank, q.q_cell_abs, q.qstart,
                       qpad, q.qrec, q.qcell, q.qpad, b.scalars);
  return hipGetLastError();
}

}  // namespace fk
