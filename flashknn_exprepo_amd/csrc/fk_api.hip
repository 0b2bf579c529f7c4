// C ABI of the FlashKNN gfx950 library (declared in include/flashknn.h).
//
// Host-side runtime: a grid handle owns the device tables of one build and
// grows them on demand (or uses a reserved capacity for graph-capturable,
// synchronisation-free builds).  Every entry point validates its arguments,
// enqueues kernels on the caller's stream and returns an fk_status; a
// message for the last failure is kept per thread (fk_last_error).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "fk_internal.h"

using namespace fk;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int ok() {
  g_err.clear();
  return FK_OK;
}

#define FK_HIP(expr)                                                        \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(FK_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define FK_CHECK(code)      \
  do {                      \
    int c_ = (code);        \
    if (c_ != FK_OK) return c_; \
  } while (0)

// Dense cells per build (sync mode). Linear cell ids are int32 and each box is
// below 2^30 cells, so a total of 2^30 keeps every id, and every scan entry
// (cells + 3), below INT32_MAX; at the cap the four cell tables take 16 GiB.
constexpr int64_t kHardCellCap = 1ll << 30;

// Balanced first pass: self-query launches from this size (below, groups
// widen to fill the GPU instead) up to kPlanMaxTilesHost super-tiles.
constexpr int32_t kPlanMinQueries = 65536;
constexpr int64_t kPlanMaxTilesHost = 8192;

// Voxel keys hold the cloud id in bits 48..62 (bit 63 clear marks a used
// hash slot): at most 2^15 - 1 clouds, the batch limit of the reference's
// xyz2key (FlashKNN/functions/z_order.py:72-91).
constexpr int32_t kMaxVoxelClouds = 32767;

inline hipStream_t S(void *s) { return (hipStream_t)s; }

}  // namespace

struct fk_grid {
  int device = -1;
  int32_t *scalars = nullptr;
  int32_t *h_pinned = nullptr;
  // per cloud
  int64_t cap_clouds = 0;
  // clouds every per-cloud buffer holds (their capacities differ: the boxes
  // take box_words(B) words, the others one entry per cloud)
  int64_t max_clouds = 0;
  int32_t *bbox = nullptr;
  int64_t cap_bbox = 0;
  CloudInfo *clouds = nullptr;
  int32_t *item_prefix = nullptr;
  int64_t cap_prefix = 0;
  unsigned long long *cloud_acc = nullptr;
  int64_t cap_acc = 0;
  // per point
  int64_t cap_points = 0, cap_p1 = 0, cap_p2 = 0, cap_p3 = 0, cap_p4 = 0;
  int4 *pt_cell = nullptr;
  int32_t *pt_lin = nullptr, *pt_rank = nullptr;
  float4 *srec = nullptr;
  int4 *scell = nullptr;
  int64_t cap_scell = 0;
  // column layout: radix sort buffers
  uint64_t *ckeys0 = nullptr, *ckeys1 = nullptr;
  int32_t *cperm0 = nullptr, *cperm1 = nullptr;
  uint8_t *sort_tmp = nullptr;
  int64_t cap_ck0 = 0, cap_ck1 = 0, cap_cp0 = 0, cap_cp1 = 0, cap_tmp = 0;
  // per cell
  int64_t cap_cells = 0, cap_c2 = 0;
  int32_t *counts = nullptr, *cell_start = nullptr;
  int64_t cap_qcells = 0, cap_qc2 = 0;
  int32_t *qcounts = nullptr, *qstart = nullptr;
  int64_t cap_part = 0;
  int32_t *scan_part = nullptr;
  // per query
  int64_t cap_q = 0, cap_q1 = 0, cap_q2 = 0, cap_q3 = 0, cap_q4 = 0, cap_q5 = 0;
  float4 *qrec = nullptr;
  int4 *qcell = nullptr, *q_cell_abs = nullptr;
  int32_t *qpad = nullptr, *q_lin = nullptr, *q_rank = nullptr;
  // exact mode
  int64_t cap_pend = 0, cap_pkeys = 0;
  int32_t *pend_q = nullptr;
  uint64_t *pend_keys = nullptr;
  // host state of the last build
  int built = 0, from_xyz = 0, num_down = 0, reserved = 0;
  int columns = 0;  // last build used the column layout
  // query counter sets: the set of the last plain first pass, which sets are
  // known to be zero, and the set fk_query_stats reports
  int qset = 0, stat_set = 0;
  // exact queries of the pass the statistics cover (certified = these minus
  // the pending count), 0 for the 27-cell mode; the same for the batched set
  int64_t stat_nq = 0, batched_nq = 0;
  bool qclean[2] = {true, true};
  int64_t n_points = 0, n_cells = -1, n_queries = -1;
  // upper bound of the last build's table entries (its exact count after a
  // synchronising build, the capacity it was held to after a no-sync one)
  int64_t build_cap = 0;
  int32_t n_clouds = 0;
  // captured batched step
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // batched first pass: the K classes above the first run on these streams
  // (default priority), forked from and joined back to the caller's stream
  hipStream_t aux[kBatchedClasses - 1] = {};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[kBatchedClasses - 1] = {};
  // balanced first pass of self queries: per-wave plan, super-tile costs,
  // and what the plan was computed for
  int2 *plan = nullptr;
  int64_t cap_plan = 0;
  int32_t *tile_cost = nullptr;
  int64_t cap_tcost = 0;
  uint64_t plan_generation = ~0ull;
  int32_t plan_kc = 0, plan_R = 0, plan_capacity = 0;
  // bumped by every build and by every reallocation of a device table: a
  // graph captured at another generation holds stale pointers and sizes
  uint64_t generation = 0, graph_generation = 0;
};

namespace {

// Grow a device array to at least `need` elements (25% + 64 slack); the old
// contents are not kept.  A reallocation moves the handle to a new
// generation (captured graphs hold the old pointers).
template <class T>
int grow(fk_grid *g, T *&p, int64_t &cap, int64_t need, const char *what) {
  if (need <= cap && p) return FK_OK;
  int64_t n = need + need / 4 + 64;
  ++g->generation;
  if (p) {
    FK_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
  }
  hipError_t e = hipMalloc((void **)&p, sizeof(T) * (size_t)n);
  if (e != hipSuccess)
    return fail(FK_ERR_HIP, "hipMalloc of %s (%lld elements) failed: %s", what,
                (long long)n, hipGetErrorString(e));
  cap = n;
  return FK_OK;
}

// Exactly `n` elements (reallocated when the capacity differs).
template <class T>
int realloc_exact(fk_grid *g, T *&p, int64_t &cap, int64_t n, const char *what) {
  if (p && cap == n) return FK_OK;
  ++g->generation;
  if (p) {
    FK_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
  }
  hipError_t e = hipMalloc((void **)&p, sizeof(T) * (size_t)n);
  if (e != hipSuccess)
    return fail(FK_ERR_HIP, "hipMalloc of %s (%lld elements) failed: %s", what,
                (long long)n, hipGetErrorString(e));
  cap = n;
  return FK_OK;
}

int bind_device(fk_grid *g) {
  int dev = 0;
  FK_HIP(hipGetDevice(&dev));
  if (g->device < 0) {
    g->device = dev;
    FK_HIP(hipMalloc((void **)&g->scalars, sizeof(int32_t) * kNumScalars));
    FK_HIP(hipMemset(g->scalars, 0, sizeof(int32_t) * kNumScalars));
    FK_HIP(hipHostMalloc((void **)&g->h_pinned, sizeof(int32_t) * kNumScalars,
                         hipHostMallocDefault));
  } else if (dev != g->device) {
    return fail(FK_ERR_STATE, "grid handle belongs to device %d, current is %d",
                g->device, dev);
  }
  return FK_OK;
}

int ensure_clouds(fk_grid *g, int64_t B) {
  FK_CHECK(grow(g, g->bbox, g->cap_bbox, box_words(B), "bbox"));
  FK_CHECK(grow(g, g->clouds, g->cap_clouds, B, "clouds"));
  FK_CHECK(grow(g, g->item_prefix, g->cap_prefix, kBatchedClasses * (B + 1),
                "item prefix"));
  FK_CHECK(grow(g, g->cloud_acc, g->cap_acc, B, "per-cloud sums"));
  if (B > g->max_clouds) g->max_clouds = B;
  return FK_OK;
}

int ensure_points(fk_grid *g, int64_t n) {
  const int64_t m = n < 1 ? 1 : n;
  FK_CHECK(grow(g, g->pt_cell, g->cap_points, m, "point cells"));
  FK_CHECK(grow(g, g->pt_lin, g->cap_p1, m, "point cell ids"));
  FK_CHECK(grow(g, g->pt_rank, g->cap_p2, m, "point ranks"));
  FK_CHECK(grow(g, g->srec, g->cap_p3, m, "sorted records"));
  FK_CHECK(grow(g, g->scell, g->cap_scell, m, "sorted cells"));
  return FK_OK;
}

// The cell tables of a build (counts, starts) and of an attach (query
// counts, starts) are sized together from one entry capacity c: counts
// c + 1, starts c + 2 (query tables c + 2 and c + 3, one extra bucket), so
// that a no-sync build held to cap_cells - 1 entries, or a no-sync attach
// held to cap_qcells - 1, always fits its start table.
int ensure_cells(fk_grid *g, int64_t cells) {
  if (g->counts && g->cell_start && cells + 1 <= g->cap_cells) return FK_OK;
  const int64_t c = cells + cells / 4 + 64;
  FK_CHECK(realloc_exact(g, g->counts, g->cap_cells, c + 1, "cell counts"));
  FK_CHECK(realloc_exact(g, g->cell_start, g->cap_c2, c + 2, "cell starts"));
  FK_CHECK(grow(g, g->scan_part, g->cap_part, scan_parts_needed(c + 2), "scan"));
  return FK_OK;
}

int ensure_qcells(fk_grid *g, int64_t cells) {
  if (g->qcounts && g->qstart && cells + 2 <= g->cap_qcells) return FK_OK;
  const int64_t c = cells + cells / 4 + 64;
  FK_CHECK(realloc_exact(g, g->qcounts, g->cap_qcells, c + 2, "query cell counts"));
  FK_CHECK(realloc_exact(g, g->qstart, g->cap_qc2, c + 3, "query cell starts"));
  FK_CHECK(grow(g, g->scan_part, g->cap_part, scan_parts_needed(c + 3), "scan"));
  return FK_OK;
}

int ensure_queries(fk_grid *g, int64_t nq) {
  const int64_t m = nq < 1 ? 1 : nq;
  FK_CHECK(grow(g, g->qrec, g->cap_q, m, "query records"));
  FK_CHECK(grow(g, g->qcell, g->cap_q1, m, "query cells"));
  FK_CHECK(grow(g, g->q_cell_abs, g->cap_q2, m, "query cells (input order)"));
  FK_CHECK(grow(g, g->qpad, g->cap_q3, m, "query pads"));
  FK_CHECK(grow(g, g->q_lin, g->cap_q4, m, "query cell ids"));
  FK_CHECK(grow(g, g->q_rank, g->cap_q5, m, "query ranks"));
  return FK_OK;
}

int ensure_pending(fk_grid *g, int64_t nq, int k) {
  const int64_t m = nq < 1 ? 1 : nq;
  FK_CHECK(grow(g, g->pend_q, g->cap_pend, m, "pending queries"));
  FK_CHECK(grow(g, g->pend_keys, g->cap_pkeys, m * k, "pending keys"));
  return FK_OK;
}

BuildBuffers build_buffers(fk_grid *g) {
  BuildBuffers b;
  b.scalars = g->scalars;
  b.bbox = g->bbox;
  b.clouds = g->clouds;
  b.pt_cell = g->pt_cell;
  b.pt_lin = g->pt_lin;
  b.pt_rank = g->pt_rank;
  b.counts = g->counts;
  b.cell_start = g->cell_start;
  b.scan_part = g->scan_part;
  b.srec = g->srec;
  b.scell = g->scell;
  b.cap_cells = g->cap_cells - 1;
  b.ckeys[0] = g->ckeys0;
  b.ckeys[1] = g->ckeys1;
  b.cperm[0] = g->cperm0;
  b.cperm[1] = g->cperm1;
  b.sort_tmp = g->sort_tmp;
  b.sort_tmp_bytes = (size_t)g->cap_tmp;
  return b;
}

__global__ void k_reset_scalars(int32_t *s) {
  if (threadIdx.x < kNumScalars) s[threadIdx.x] = 0;
}

int read_scalars(fk_grid *g, hipStream_t s) {
  FK_HIP(hipMemcpyAsync(g->h_pinned, g->scalars, sizeof(int32_t) * kNumScalars,
                        hipMemcpyDeviceToHost, s));
  FK_HIP(hipStreamSynchronize(s));
  return FK_OK;
}

int status_message(int st) {
  switch (st) {
    case FK_ERR_DOMAIN:
      return fail(st, "a batch id or cell coordinate is outside the supported "
                      "domain (batch in [0, num_clouds), |cell| < 2^30, finite "
                      "coordinates)");
    case FK_ERR_GRID_TOO_LARGE:
      return fail(st, "the dense per-cloud cell boxes exceed the cell capacity; "
                      "use a larger cell size (or reserve more cells)");
    case FK_ERR_K_RANGE:
      return fail(st, "a per-cloud k is outside [1, params->k] of the batched "
                      "launch");
    case FK_ERR_STATE:
      return fail(st, "the pending-query counter was not reset before a "
                      "first pass (passes mask without bit 0)");
    default:
      return fail(st, "device status %d", st);
  }
}

// Common tail of both builds: layout, optional size read-back, sort.
int finish_build(fk_grid *g, const float *xyz, const int64_t *grid, int64_t n,
                 int32_t B, const int32_t *offsets, const float *inv_cell,
                 int32_t flags, hipStream_t s) {
  const bool nosync = flags & FK_BUILD_NO_SYNC;
  BuildBuffers bb = build_buffers(g);
  if (nosync) {
    // A no-sync build stays within the reserved query tables too, so that a
    // no-sync attach always fits (a synchronising build may have grown the
    // cell tables past the reserve).
    if (g->qcounts && g->cap_qcells - 1 < bb.cap_cells) bb.cap_cells = g->cap_qcells - 1;
  } else {
    bb.cap_cells = kHardCellCap;
  }
  g->columns = (flags & FK_BUILD_COLUMNS) ? 1 : 0;
  FK_HIP(launch_layout(bb, B, offsets, inv_cell, g->columns, s));
  int64_t hint;
  if (nosync) {
    hint = bb.cap_cells;
    g->n_cells = -1;
  } else {
    FK_CHECK(read_scalars(g, s));
    if (g->h_pinned[kStatus] == FK_ERR_GRID_TOO_LARGE && !g->columns) {
      // Too many cells for one table entry each: one entry per (x, y)
      // column, points sorted by z inside the column.
      FK_HIP(launch_fill_i32(g->scalars + kStatus, 0, 1, s));
      FK_HIP(launch_layout(bb, B, offsets, inv_cell, 1, s));
      FK_CHECK(read_scalars(g, s));
      g->columns = 1;
    }
    if (g->h_pinned[kStatus] != 0) return status_message(g->h_pinned[kStatus]);
    hint = g->h_pinned[kTotalCells];
    FK_CHECK(ensure_cells(g, hint));
    g->n_cells = hint;
    bb = build_buffers(g);
  }
  if (g->columns) {
    auto bits = [](uint64_t v) {  // bits to hold 0..v
      int b = 0;
      while (v >> b) ++b;
      return b;
    };
    const int depth = g->h_pinned[kMaxDepth];
    const int depth_bits = bits(depth > 0 ? (uint64_t)depth - 1 : 0);
    const int key_bits = std::max(1, depth_bits + bits((uint64_t)hint));
    size_t tmp = 0;
    FK_HIP(column_sort_tmp_bytes(n, key_bits, &tmp));
    const int64_t m = n < 1 ? 1 : n;
    FK_CHECK(grow(g, g->ckeys0, g->cap_ck0, m, "column keys"));
    FK_CHECK(grow(g, g->ckeys1, g->cap_ck1, m, "column keys"));
    FK_CHECK(grow(g, g->cperm0, g->cap_cp0, m, "column order"));
    FK_CHECK(grow(g, g->cperm1, g->cap_cp1, m, "column order"));
    FK_CHECK(grow(g, g->sort_tmp, g->cap_tmp, (int64_t)tmp + 1, "column sort"));
    bb = build_buffers(g);
    FK_HIP(launch_sort_columns(bb, xyz, grid, n, hint, depth_bits, key_bits, s));
  } else {
    FK_HIP(launch_sort(bb, xyz, grid, n, hint, s));
  }
  g->build_cap = hint;
  g->built = 1;
  ++g->generation;  // new tables and sizes: a captured step is stale
  g->n_points = n;
  g->n_clouds = B;
  g->n_queries = -1;
  return ok();
}

int check_stream_device(fk_grid *g) {
  if (!g) return fail(FK_ERR_INVALID_ARG, "grid handle is NULL");
  return bind_device(g);
}

}  // namespace

extern "C" {

const char *fk_status_string(int status) {
  switch (status) {
    case FK_OK: return "ok";
    case FK_ERR_INVALID_ARG: return "invalid argument";
    case FK_ERR_K_RANGE: return "k out of range";
    case FK_ERR_GRID_TOO_LARGE: return "grid too large";
    case FK_ERR_DOMAIN: return "input outside domain";
    case FK_ERR_HIP: return "HIP runtime error";
    case FK_ERR_CAPACITY: return "reserved capacity exceeded";
    case FK_ERR_STATE: return "invalid handle state";
    default: return "unknown status";
  }
}

const char *fk_last_error(void) { return g_err.c_str(); }

int fk_abi_version(void) { return FK_ABI_VERSION; }

int fk_morton_encode(const int64_t *grid_coord, const int64_t *batch_idx,
                     int64_t n, int64_t *keys, void *stream) {
  if (n < 0 || (n > 0 && (!grid_coord || !keys)))
    return fail(FK_ERR_INVALID_ARG, "fk_morton_encode: bad arguments");
  FK_HIP(launch_morton_encode(grid_coord, batch_idx, n, keys, S(stream)));
  return ok();
}

int fk_morton_neighbor_keys(const int64_t *grid_down,
                            const int64_t *parent_batch, int64_t num_parents,
                            int64_t *keys, void *stream) {
  if (num_parents < 0 || (num_parents > 0 && (!grid_down || !keys)))
    return fail(FK_ERR_INVALID_ARG, "fk_morton_neighbor_keys: bad arguments");
  FK_HIP(launch_morton_neighbors(grid_down, parent_batch, num_parents, keys,
                                 S(stream)));
  return ok();
}

int fk_grid_create(fk_grid **out) {
  if (!out) return fail(FK_ERR_INVALID_ARG, "fk_grid_create: out is NULL");
  *out = new fk_grid();
  return ok();
}

int fk_grid_destroy(fk_grid *g) {
  if (!g) return ok();
  // Teardown keeps going past individual failures; the first is reported.
  hipError_t first = hipSuccess;
  auto note = [&first](hipError_t e) {
    if (first == hipSuccess) first = e;
  };
  if (g->gexec) note(hipGraphExecDestroy(g->gexec));
  if (g->graph) note(hipGraphDestroy(g->graph));
  for (int i = 0; i < kBatchedClasses - 1; ++i) {
    if (g->aux[i]) note(hipStreamDestroy(g->aux[i]));
    if (g->ev_join[i]) note(hipEventDestroy(g->ev_join[i]));
  }
  if (g->ev_fork) note(hipEventDestroy(g->ev_fork));
  void *bufs[] = {g->scalars, g->bbox, g->clouds, g->item_prefix, g->cloud_acc,
                  g->pt_cell,
                  g->pt_lin, g->pt_rank, g->srec, g->scell, g->counts,
                  g->cell_start, g->qcounts, g->qstart, g->scan_part, g->qrec,
                  g->qcell, g->q_cell_abs, g->qpad, g->q_lin, g->q_rank,
                  g->pend_q, g->pend_keys, g->ckeys0, g->ckeys1, g->cperm0,
                  g->cperm1, g->sort_tmp, g->plan, g->tile_cost};
  for (void *p : bufs)
    if (p) note(hipFree(p));
  if (g->h_pinned) note(hipHostFree(g->h_pinned));
  delete g;
  if (first != hipSuccess)
    return fail(FK_ERR_HIP, "fk_grid_destroy: %s", hipGetErrorString(first));
  return ok();
}

int fk_grid_reserve(fk_grid *g, int64_t max_points, int64_t max_cells,
                    int32_t max_clouds, int64_t max_queries) {
  FK_CHECK(check_stream_device(g));
  if (max_points < 0 || max_cells < 0 || max_clouds < 1 || max_queries < 0 ||
      max_points > INT32_MAX || max_cells > kHardCellCap)
    return fail(FK_ERR_INVALID_ARG, "fk_grid_reserve: bad sizes");
  // A table reallocated here loses its contents: a build whose tables moved
  // is gone (queries are refused until the next build), and so are attached
  // queries whose tables moved.  Checked on the failure paths too.
  const uint64_t gen_build = g->generation;
  int st = ensure_clouds(g, max_clouds);
  if (st == FK_OK) st = ensure_points(g, max_points);
  if (st == FK_OK) st = ensure_cells(g, max_cells);
  if (g->generation != gen_build) {
    g->built = 0;
    g->n_queries = -1;
  }
  if (st != FK_OK) return st;
  const uint64_t gen_queries = g->generation;
  const int64_t nq = max_queries > max_points ? max_queries : max_points;
  st = ensure_qcells(g, max_cells);
  if (st == FK_OK) st = ensure_queries(g, nq);
  if (g->generation != gen_queries) g->n_queries = -1;
  if (st != FK_OK) return st;
  FK_CHECK(ensure_pending(g, nq, kMaxLaneK));
  g->reserved = 1;
  return ok();
}

int fk_grid_build_cells(fk_grid *g, const float *xyz, const int64_t *grid_coord,
                        const int64_t *batch_idx, int64_t n, int32_t num_clouds,
                        int32_t num_down, int32_t flags, void *stream) {
  FK_CHECK(check_stream_device(g));
  if (n < 0 || n > INT32_MAX || num_clouds < 1 || num_down < 0 || num_down > 30 ||
      (n > 0 && !grid_coord))
    return fail(FK_ERR_INVALID_ARG,
                "fk_grid_build_cells: n in [0, 2^31), num_clouds >= 1, "
                "num_down in [0, 30] and grid_coord required");
  // Until this build succeeds the handle holds no usable build: a failure
  // below may already have reallocated (or, on a failed hipMalloc, freed)
  // tables that the previous build's sizes describe.
  g->built = 0;
  g->n_queries = -1;
  ++g->generation;
  hipStream_t s = S(stream);
  const bool nosync = flags & FK_BUILD_NO_SYNC;
  if (nosync && (flags & FK_BUILD_COLUMNS))
    return fail(FK_ERR_INVALID_ARG, "FK_BUILD_COLUMNS needs a synchronising build");
  if (nosync) {
    if (n > g->cap_points || num_clouds > g->max_clouds || !g->counts)
      return fail(FK_ERR_CAPACITY, "FK_BUILD_NO_SYNC needs fk_grid_reserve "
                                   "of at least n points and num_clouds clouds");
  } else {
    FK_CHECK(ensure_clouds(g, num_clouds));
    FK_CHECK(ensure_points(g, n));
    if (!g->counts) FK_CHECK(ensure_cells(g, 1024));
  }
  hipLaunchKernelGGL(k_reset_scalars, dim3(1), dim3(64), 0, s, g->scalars);
  g->qclean[0] = g->qclean[1] = true;
  FK_HIP(launch_build_cells(build_buffers(g), xyz, grid_coord, batch_idx, n,
                            num_clouds, num_down, s));
  g->from_xyz = 0;
  g->num_down = num_down;
  return finish_build(g, xyz, grid_coord, n, num_clouds, nullptr, nullptr, flags,
                      s);
}

int fk_grid_build_xyz(fk_grid *g, const float *xyz, const int32_t *offsets,
                      int32_t num_clouds, int64_t n, const float *inv_cell,
                      int32_t flags, void *stream) {
  FK_CHECK(check_stream_device(g));
  if (n < 0 || n > INT32_MAX || num_clouds < 1 || !offsets || !inv_cell ||
      (n > 0 && !xyz))
    return fail(FK_ERR_INVALID_ARG,
                "fk_grid_build_xyz: n in [0, 2^31), num_clouds >= 1, xyz, "
                "offsets and inv_cell required");
  // Until this build succeeds the handle holds no usable build: a failure
  // below may already have reallocated (or, on a failed hipMalloc, freed)
  // tables that the previous build's sizes describe.
  g->built = 0;
  g->n_queries = -1;
  ++g->generation;
  hipStream_t s = S(stream);
  const bool nosync = flags & FK_BUILD_NO_SYNC;
  if (nosync && (flags & FK_BUILD_COLUMNS))
    return fail(FK_ERR_INVALID_ARG, "FK_BUILD_COLUMNS needs a synchronising build");
  if (nosync) {
    if (n > g->cap_points || num_clouds > g->max_clouds || !g->counts)
      return fail(FK_ERR_CAPACITY, "FK_BUILD_NO_SYNC needs fk_grid_reserve "
                                   "of at least n points and num_clouds clouds");
  } else {
    FK_CHECK(ensure_clouds(g, num_clouds));
    FK_CHECK(ensure_points(g, n));
    if (!g->counts) FK_CHECK(ensure_cells(g, 1024));
  }
  hipLaunchKernelGGL(k_reset_scalars, dim3(1), dim3(64), 0, s, g->scalars);
  g->qclean[0] = g->qclean[1] = true;
  FK_HIP(launch_build_xyz_cells(build_buffers(g), xyz, offsets, num_clouds, n,
                                inv_cell, s));
  g->from_xyz = 1;
  g->num_down = 0;
  return finish_build(g, xyz, nullptr, n, num_clouds, offsets, inv_cell, flags,
                      s);
}

int fk_grid_attach_queries(fk_grid *g, const float *qxyz, const int64_t *qcell,
                           const int64_t *qbatch, const int32_t *qoffsets,
                           int64_t nq, const int32_t *qpad, int32_t flags,
                           void *stream) {
  FK_CHECK(check_stream_device(g));
  if (!g->built) return fail(FK_ERR_STATE, "fk_grid_attach_queries: grid not built");
  if (nq < 0 || nq > INT32_MAX || (nq > 0 && !qxyz))
    return fail(FK_ERR_INVALID_ARG, "fk_grid_attach_queries: bad query arrays");
  if (g->from_xyz) {
    if (qcell || !qoffsets)
      return fail(FK_ERR_INVALID_ARG, "grids built from xyz take query offsets "
                                      "and no query cells");
  } else if (nq > 0 && !qcell) {
    return fail(FK_ERR_INVALID_ARG, "grids built from cells need query cells");
  }
  hipStream_t s = S(stream);
  const bool nosync = flags & FK_BUILD_NO_SYNC;
  g->n_queries = -1;  // set again once the query tables describe this list
  int64_t hint = g->n_cells;
  if (nosync || hint < 0) {
    // the query tables hold the build's table entries + 1 (one bucket for
    // queries of clouds without cells)
    if (nq > g->cap_q || !g->qcounts || g->build_cap > g->cap_qcells - 1)
      return fail(FK_ERR_CAPACITY, "FK_BUILD_NO_SYNC needs reserved query capacity");
    hint = g->build_cap;
  } else {
    FK_CHECK(ensure_queries(g, nq));
    FK_CHECK(ensure_qcells(g, hint));
  }
  QueryBuffers q;
  q.qrec = g->qrec;
  q.qcell = g->qcell;
  q.qpad = g->qpad;
  q.q_lin = g->q_lin;
  q.q_rank = g->q_rank;
  q.q_cell_abs = g->q_cell_abs;
  q.qcounts = g->qcounts;
  q.qstart = g->qstart;
  FK_HIP(launch_attach(build_buffers(g), q, qxyz, qcell, qbatch, qoffsets, nq,
                       qpad, g->n_clouds, g->num_down, g->from_xyz, hint, s));
  g->n_queries = nq;
  if (!nosync) {
    FK_CHECK(read_scalars(g, s));
    if (g->h_pinned[kStatus] != 0) {
      g->n_queries = -1;
      return status_message(g->h_pinned[kStatus]);
    }
  }
  return ok();
}

int fk_grid_status(fk_grid *g, void *stream) {
  FK_CHECK(check_stream_device(g));
  FK_CHECK(read_scalars(g, S(stream)));
  const int st = g->h_pinned[kStatus];
  if (st != 0) return status_message(st);
  if (g->built) g->n_cells = g->h_pinned[kTotalCells];  // known after a no-sync build
  return ok();
}

int fk_grid_info(fk_grid *g, int64_t *num_points, int32_t *num_clouds,
                 int64_t *num_cells, int64_t *num_queries) {
  if (!g) return fail(FK_ERR_INVALID_ARG, "grid handle is NULL");
  if (num_points) *num_points = g->built ? g->n_points : -1;
  if (num_clouds) *num_clouds = g->built ? g->n_clouds : -1;
  if (num_cells) *num_cells = g->built ? g->n_cells : -1;
  if (num_queries) *num_queries = g->n_queries;
  return ok();
}

int fk_grid_tables(fk_grid *g, const void **sorted_records,
                   const int32_t **cell_start) {
  if (!g || !g->built) return fail(FK_ERR_STATE, "grid not built");
  if (sorted_records) *sorted_records = g->srec;
  if (cell_start) *cell_start = g->cell_start;
  return ok();
}

static int make_args(fk_grid *g, const fk_query_params *p, void *out_idx,
                     float *out_dist, QueryArgs &a) {
  if (!g || !p) return fail(FK_ERR_INVALID_ARG, "fk_query: NULL handle or params");
  if (!g->built) return fail(FK_ERR_STATE, "fk_query: grid not built");
  if (p->k < 1 || p->k > FK_MAX_K)
    return fail(FK_ERR_K_RANGE, "fk_query: k = %d outside [1, %d]", p->k, FK_MAX_K);
  if (p->mode != FK_MODE_CELLS && p->mode != FK_MODE_EXACT)
    return fail(FK_ERR_INVALID_ARG, "fk_query: unknown mode %d", p->mode);
  if (p->mode == FK_MODE_EXACT && !g->from_xyz)
    return fail(FK_ERR_INVALID_ARG,
                "FK_MODE_EXACT needs a grid built by fk_grid_build_xyz");
  if (p->radius_cells < 1 || p->radius_cells > 1024)
    return fail(FK_ERR_INVALID_ARG, "fk_query: radius_cells must be in [1, 1024]");
  if (p->pad_mode < FK_PAD_SELF || p->pad_mode > FK_PAD_PER_QUERY)
    return fail(FK_ERR_INVALID_ARG, "fk_query: unknown pad_mode");
  if (p->dist_mode < FK_DIST_NONE || p->dist_mode > FK_DIST_SQRT)
    return fail(FK_ERR_INVALID_ARG, "fk_query: unknown dist_mode");
  if (p->query_set != FK_QUERIES_SELF && p->query_set != FK_QUERIES_ATTACHED)
    return fail(FK_ERR_INVALID_ARG, "fk_query: unknown query_set");
  if (p->query_set == FK_QUERIES_ATTACHED && g->n_queries < 0)
    return fail(FK_ERR_STATE, "fk_query: no queries attached");
  if (p->pad_mode == FK_PAD_PER_QUERY && p->query_set != FK_QUERIES_ATTACHED)
    return fail(FK_ERR_INVALID_ARG, "FK_PAD_PER_QUERY needs attached queries");
  if (p->pad_mode == FK_PAD_SELF && p->query_set != FK_QUERIES_SELF)
    return fail(FK_ERR_INVALID_ARG, "FK_PAD_SELF is for self queries");
  const bool self = p->query_set == FK_QUERIES_SELF;
  const int64_t nq = self ? g->n_points : g->n_queries;
  // an empty query set may pass NULL outputs (a 0-element tensor has none)
  if (!out_idx && nq > 0) return fail(FK_ERR_INVALID_ARG, "fk_query: out_idx is NULL");
  if (p->dist_mode != FK_DIST_NONE && !out_dist && nq > 0)
    return fail(FK_ERR_INVALID_ARG, "fk_query: out_dist is NULL");
  if (std::isnan(p->cut_r2))
    return fail(FK_ERR_INVALID_ARG, "fk_query: cut_r2 is NaN");
  if ((p->flags & FK_QUERY_BATCH_SMALL) && (p->flags & FK_QUERY_BATCH_LARGE))
    return fail(FK_ERR_INVALID_ARG, "fk_query: FK_QUERY_BATCH_SMALL and _LARGE together");
  memset(&a, 0, sizeof(a));
  a.srec = g->srec;
  a.scell = g->scell;
  a.cell_start = g->cell_start;
  a.columns = g->columns;
  a.clouds = g->clouds;
  a.qrec = self ? g->srec : g->qrec;
  a.qcell = self ? g->scell : g->qcell;
  a.qpad = self ? nullptr : g->qpad;
  a.nq = (int32_t)(self ? g->n_points : g->n_queries);
  a.k = p->k;
  a.R = p->radius_cells;
  // d < cut with cut <= 0 holds for no candidate: a cut of +0 says the same
  a.cut_r2 = p->cut_r2 > 0.f ? p->cut_r2 : 0.f;
  a.pad_mode = p->pad_mode;
  a.dist_mode = p->dist_mode;
  a.pad_dist = p->pad_dist;
  a.idx64 = p->idx_64 ? 1 : 0;
  a.out_idx = out_idx;
  a.out_d = p->dist_mode == FK_DIST_NONE ? nullptr : out_dist;
  a.exact = p->mode == FK_MODE_EXACT;
  a.passes = p->passes ? (p->passes & 7) : 7;
  a.prune = (p->flags & FK_QUERY_PRUNE) ? 1 : 0;
  a.one_lane = (p->flags & FK_QUERY_ONE_LANE) ? 1 : 0;
  a.batch_sel = (p->flags & FK_QUERY_BATCH_SMALL) ? 1 : (p->flags & FK_QUERY_BATCH_LARGE) ? 2 : 0;
  a.scalars = g->scalars;
  a.qcnt = g->scalars + kCounters + kCntStride * g->qset;
  a.qcnt_next = nullptr;
  a.pend_class = -1;
  return FK_OK;
}

// The balanced first pass (launch_plan) for self queries of K classes
// 8-32 on launches large enough not to widen their lane groups and small
// enough for one resident round (at most kPlanTile queries per resident
// wave); the plan is computed once per build and kept.
static int attach_plan(fk_grid *g, const fk_query_params *p, QueryArgs &a,
                       hipStream_t s) {
  a.plan = nullptr;
  a.plan_waves = 0;
  if (p->query_set != FK_QUERIES_SELF || a.k > 32 || a.one_lane ||
      (p->flags & FK_QUERY_UNIFORM) || !(a.passes & 2) || a.nq < kPlanMinQueries)
    return FK_OK;
  const int kc = kclass_of(a.k);
  int32_t cap = 0;
  FK_HIP(planned_capacity(kc, a.batch_sel, &cap));
  const int64_t ntiles = ((int64_t)a.nq + kPlanTile - 1) / kPlanTile;
  if (ntiles > cap || ntiles > kPlanMaxTilesHost) return FK_OK;
  FK_CHECK(grow(g, g->plan, g->cap_plan, cap, "first-pass plan"));
  FK_CHECK(grow(g, g->tile_cost, g->cap_tcost, ntiles, "super-tile costs"));
  if (g->plan_generation != g->generation || g->plan_kc != kc || g->plan_R != a.R ||
      g->plan_capacity != cap) {
    FK_HIP(launch_plan(a, cap, g->tile_cost, g->plan, s));
    g->plan_generation = g->generation;
    g->plan_kc = kc;
    g->plan_R = a.R;
    g->plan_capacity = cap;
  }
  a.plan = g->plan;
  a.plan_waves = cap;
  return FK_OK;
}

int fk_query(fk_grid *g, const fk_query_params *p, void *out_idx,
             float *out_dist, void *stream) {
  FK_CHECK(check_stream_device(g));
  QueryArgs a;
  FK_CHECK(make_args(g, p, out_idx, out_dist, a));
  if (a.exact && a.k <= kMaxLaneK) {
    // grows only when the reserve (or the last query) was smaller
    FK_CHECK(ensure_pending(g, a.nq, a.k));
    a.pend_q = g->pend_q;
    a.pend_keys = g->pend_keys;
  }
  // A first pass starts a new counter set (alternating); it needs a reset
  // launch only when that set is not known to be zero (the previous first
  // pass clears it).  An explicit bit 0 in params->passes always resets.
  const bool first = a.passes & 2;
  const int set = first ? (g->qset ^ 1) : g->qset;
  const bool reset = (p->passes && (a.passes & 1)) || (first && !g->qclean[set]);
  a.qcnt = g->scalars + kCounters + kCntStride * set;
  a.qcnt_next = first && a.nq > 0 ? g->scalars + kCounters + kCntStride * (set ^ 1) : nullptr;
  a.passes = (a.passes & ~1) | (reset ? 1 : 0);
  FK_CHECK(attach_plan(g, p, a, S(stream)));
  FK_HIP(launch_query(a, S(stream)));
  g->qset = set;
  g->stat_set = set;
  if (first) g->stat_nq = a.exact ? a.nq : 0;
  if (reset) g->qclean[set] = true;
  if (first) {
    g->qclean[set] = false;
    if (a.qcnt_next) g->qclean[set ^ 1] = true;
  }
  return ok();
}

int fk_query_stats(fk_grid *g, int64_t *first_pass_certified, int64_t *expanded,
                   void *stream) {
  FK_CHECK(check_stream_device(g));
  FK_CHECK(read_scalars(g, S(stream)));
  // a batched launch keeps one set per K class
  const int sets = g->stat_set == kBatchedSet ? kBatchedClasses : 1;
  int64_t listed = 0, exp = 0;
  for (int i = 0; i < sets; ++i) {
    const int32_t *c = g->h_pinned + kCounters + kCntStride * (g->stat_set + i);
    const int64_t inl = c[kCntInline] < kInlineBudget ? c[kCntInline] : kInlineBudget;
    listed += (int64_t)c[kCntPending] + inl;
    exp += c[kCntExpanded];
  }
  const int64_t pending = listed < g->stat_nq ? listed : g->stat_nq;
  if (first_pass_certified) *first_pass_certified = g->stat_nq - pending;
  if (expanded) *expanded = exp;
  return ok();
}

int fk_candidate_count(fk_grid *g, int32_t radius_cells, int32_t query_set,
                       int64_t *total, void *stream) {
  FK_CHECK(check_stream_device(g));
  if (!total) return fail(FK_ERR_INVALID_ARG, "fk_candidate_count: total is NULL");
  fk_query_params p;
  memset(&p, 0, sizeof(p));
  p.k = 1;
  p.mode = FK_MODE_CELLS;
  p.radius_cells = radius_cells;
  p.cut_r2 = INFINITY;
  p.pad_mode = query_set == FK_QUERIES_SELF ? FK_PAD_SELF : FK_PAD_MINUS1;
  p.query_set = query_set;
  QueryArgs a;
  int dummy = 0;
  FK_CHECK(make_args(g, &p, &dummy, nullptr, a));
  hipStream_t s = S(stream);
  unsigned long long *acc = (unsigned long long *)(g->scalars + 8);
  FK_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long), s));
  FK_HIP(launch_candidate_count(a, acc, s));
  unsigned long long h = 0;
  FK_HIP(hipMemcpyAsync(&h, acc, sizeof(h), hipMemcpyDeviceToHost, s));
  FK_HIP(hipStreamSynchronize(s));
  *total = (int64_t)h;
  return ok();
}

int fk_point_candidates(fk_grid *g, int32_t radius_cells, int32_t query_set,
                        int32_t *out, void *stream) {
  FK_CHECK(check_stream_device(g));
  fk_query_params p;
  memset(&p, 0, sizeof(p));
  p.k = 1;
  p.mode = FK_MODE_CELLS;
  p.radius_cells = radius_cells;
  p.cut_r2 = INFINITY;
  p.pad_mode = query_set == FK_QUERIES_SELF ? FK_PAD_SELF : FK_PAD_MINUS1;
  p.query_set = query_set;
  QueryArgs a;
  int dummy = 0;
  FK_CHECK(make_args(g, &p, &dummy, nullptr, a));
  if (!out && a.nq > 0) return fail(FK_ERR_INVALID_ARG, "fk_point_candidates: out is NULL");
  FK_HIP(launch_point_candidates(a, out, S(stream)));
  return ok();
}

int fk_grid_auto_cell(fk_grid *g, const float *xyz, const int32_t *offsets,
                      int32_t num_clouds, int64_t n, const int32_t *k_per_cloud,
                      int32_t k, float scale, float *side_out, float *inv_out,
                      void *stream) {
  FK_CHECK(check_stream_device(g));
  if (n < 0 || num_clouds < 1 || !offsets || !inv_out || (n > 0 && !xyz) ||
      (!k_per_cloud && (k < 1 || k > FK_MAX_K)) || !(scale > 0.f))
    return fail(FK_ERR_INVALID_ARG, "fk_grid_auto_cell: bad arguments");
  if (num_clouds > g->max_clouds) {
    if (g->reserved)
      return fail(FK_ERR_CAPACITY, "fk_grid_auto_cell: more clouds than reserved");
    FK_CHECK(ensure_clouds(g, num_clouds));
  }
  FK_HIP(launch_auto_cell(g->bbox, xyz, offsets, num_clouds, n, k_per_cloud, k,
                          scale, side_out, inv_out, S(stream)));
  return ok();
}

int fk_grid_cloud_candidates(fk_grid *g, int64_t *out, void *stream) {
  FK_CHECK(check_stream_device(g));
  if (!g->built) return fail(FK_ERR_STATE, "fk_grid_cloud_candidates: grid not built");
  if (!out) return fail(FK_ERR_INVALID_ARG, "fk_grid_cloud_candidates: out is NULL");
  fk_query_params p;
  memset(&p, 0, sizeof(p));
  p.k = 1;
  p.mode = FK_MODE_CELLS;
  p.radius_cells = 1;
  p.cut_r2 = INFINITY;
  p.query_set = FK_QUERIES_SELF;
  QueryArgs a;
  int dummy = 0;
  FK_CHECK(make_args(g, &p, &dummy, nullptr, a));
  a.num_clouds = g->n_clouds;
  FK_HIP(launch_cloud_candidates(a, (unsigned long long *)out, S(stream)));
  return ok();
}

int fk_grid_adapt_cell(fk_grid *g, const int32_t *k_per_cloud, int32_t k,
                       float target_ratio, const float *target_per_cloud,
                       float exponent, float *side_io, float *inv_io,
                       void *stream) {
  FK_CHECK(check_stream_device(g));
  if (!g->built || !g->from_xyz)
    return fail(FK_ERR_STATE, "fk_grid_adapt_cell needs a grid built from xyz");
  if (!side_io || !inv_io || (!target_per_cloud && !(target_ratio > 0.f)) ||
      !(exponent > 0.f) || (!k_per_cloud && (k < 1 || k > FK_MAX_K)))
    return fail(FK_ERR_INVALID_ARG, "fk_grid_adapt_cell: bad arguments");
  fk_query_params p;
  memset(&p, 0, sizeof(p));
  p.k = 1;
  p.mode = FK_MODE_CELLS;
  p.radius_cells = 1;
  p.cut_r2 = INFINITY;
  p.query_set = FK_QUERIES_SELF;
  QueryArgs a;
  int dummy = 0;
  FK_CHECK(make_args(g, &p, &dummy, nullptr, a));
  a.num_clouds = g->n_clouds;
  FK_HIP(launch_adapt_cell(a, g->cloud_acc, k_per_cloud, k, target_ratio,
                           target_per_cloud, exponent, side_io, inv_io, S(stream)));
  return ok();
}

static int batched_args(fk_grid *g, const fk_query_params *p,
                        const int32_t *k_per_cloud, const int64_t *out_offsets,
                        void *out_idx, float *out_dist, QueryArgs &a) {
  FK_CHECK(make_args(g, p, out_idx, out_dist, a));
  if (!g->from_xyz)
    return fail(FK_ERR_INVALID_ARG, "fk_batched_query needs a grid built by "
                                    "fk_grid_build_xyz (cloud offsets)");
  if (p->query_set != FK_QUERIES_SELF)
    return fail(FK_ERR_INVALID_ARG, "fk_batched_query runs self queries");
  if (!k_per_cloud || !out_offsets)
    return fail(FK_ERR_INVALID_ARG, "fk_batched_query: k_per_cloud and "
                                    "out_offsets required");
  if (p->k > FK_MAX_BATCHED_K)
    return fail(FK_ERR_K_RANGE, "fk_batched_query: k = %d above %d", p->k,
                FK_MAX_BATCHED_K);
  if (a.exact) {
    FK_CHECK(ensure_pending(g, a.nq, a.k));
    a.pend_q = g->pend_q;
    a.pend_keys = g->pend_keys;
  }
  a.k_per_cloud = k_per_cloud;
  a.out_offsets = out_offsets;
  // class 0's counter set; class c uses set kBatchedSet + c (cleared by k_items)
  a.qcnt = g->scalars + kCounters + kCntStride * kBatchedSet;
  a.qcnt_next = nullptr;
  a.num_clouds = g->n_clouds;
  a.item_cloud = g->item_prefix;
  return FK_OK;
}

// Work items, then per K class up to max_k's one persistent first-pass
// kernel and the shell expansion of that class's pending list (the classes
// above the first forked onto the handle's auxiliary streams, so that one
// class's last tiles and expansion overlap the others' instead of leaving
// the GPU idle), joined back on s.  Graph-capturable (the fork and join are
// events).
#ifndef FK_BATCHED_FORK
#define FK_BATCHED_FORK 1
#endif
// The auxiliary streams and events (created once, on the handle's device,
// outside any capture).
static int ensure_fork(fk_grid *g) {
  if (!FK_BATCHED_FORK || g->ev_fork) return FK_OK;
  for (int i = 0; i < kBatchedClasses - 1; ++i) {
    if (!g->aux[i]) FK_HIP(hipStreamCreateWithFlags(&g->aux[i], hipStreamNonBlocking));
    if (!g->ev_join[i])
      FK_HIP(hipEventCreateWithFlags(&g->ev_join[i], hipEventDisableTiming));
  }
  FK_HIP(hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming));
  return FK_OK;
}

// One K class: its first pass, then the expansion of its own pending list
// (so that a class that finishes early expands while the others still run).
static QueryArgs class_args(const QueryArgs &a, int cls) {
  QueryArgs c = a;
  c.qcnt = a.qcnt + kCntStride * cls;
  c.pend_class = cls;
  return c;
}
static int enqueue_class(const QueryArgs &a, int cls, hipStream_t s) {
  const QueryArgs c = class_args(a, cls);
  FK_HIP(launch_batched_first(c, cls, s));
  FK_HIP(c.columns ? launch_expand_columns(c, s) : launch_expand(c, s));
  return FK_OK;
}

static int enqueue_batched(fk_grid *g, const QueryArgs &a, int32_t max_k, hipStream_t s) {
  FK_HIP(launch_items(g->clouds, g->cell_start, a.k_per_cloud, g->n_clouds, max_k,
                      g->item_prefix, g->scalars, a.qcnt, s));
  const int top = max_k <= 8 ? 0 : (max_k <= 16 ? 1 : (max_k <= 32 ? 2 : 3));
  if (FK_BATCHED_FORK && top > 0) {
    FK_HIP(hipEventRecord(g->ev_fork, s));
    for (int cls = top; cls >= 1; --cls) {
      hipStream_t as = g->aux[cls - 1];
      FK_HIP(hipStreamWaitEvent(as, g->ev_fork, 0));
      FK_CHECK(enqueue_class(a, cls, as));
      FK_HIP(hipEventRecord(g->ev_join[cls - 1], as));
    }
    FK_CHECK(enqueue_class(a, 0, s));
    for (int cls = top; cls >= 1; --cls) FK_HIP(hipStreamWaitEvent(s, g->ev_join[cls - 1], 0));
  } else {
    for (int cls = top; cls >= 0; --cls) FK_CHECK(enqueue_class(a, cls, s));
  }
  return FK_OK;
}

int fk_batched_query(fk_grid *g, const fk_query_params *p,
                     const int32_t *k_per_cloud, const int64_t *out_offsets,
                     void *out_idx, float *out_dist, void *stream) {
  FK_CHECK(check_stream_device(g));
  QueryArgs a;
  FK_CHECK(batched_args(g, p, k_per_cloud, out_offsets, out_idx, out_dist, a));
  FK_CHECK(ensure_fork(g));
  FK_CHECK(enqueue_batched(g, a, p->k, S(stream)));
  g->stat_set = kBatchedSet;
  g->batched_nq = a.exact ? a.nq : 0;
  g->stat_nq = g->batched_nq;
  return ok();
}

int fk_batched_capture(fk_grid *g, const fk_query_params *p,
                       const int32_t *k_per_cloud, const int64_t *out_offsets,
                       void *out_idx, float *out_dist, void *stream) {
  FK_CHECK(check_stream_device(g));
  if (!stream)
    return fail(FK_ERR_INVALID_ARG, "fk_batched_capture needs a non-null stream");
  QueryArgs a;
  FK_CHECK(batched_args(g, p, k_per_cloud, out_offsets, out_idx, out_dist, a));
  if (g->gexec) {
    FK_HIP(hipGraphExecDestroy(g->gexec));
    g->gexec = nullptr;
  }
  if (g->graph) {
    FK_HIP(hipGraphDestroy(g->graph));
    g->graph = nullptr;
  }
  FK_CHECK(ensure_fork(g));
  hipStream_t s = S(stream);
  FK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_batched(g, a, p->k, s);
  hipGraph_t graph = nullptr;
  hipError_t e3 = hipStreamEndCapture(s, &graph);
  if (rc != FK_OK || e3 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != FK_OK) return rc;
    return fail(FK_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e3));
  }
  g->graph = graph;
  FK_HIP(hipGraphInstantiate(&g->gexec, graph, nullptr, nullptr, 0));
  g->batched_nq = a.exact ? a.nq : 0;
  g->graph_generation = g->generation;
  return ok();
}

int fk_graph_replay(fk_grid *g, void *stream) {
  if (!g || !g->gexec) return fail(FK_ERR_STATE, "fk_graph_replay: nothing captured");
  if (g->generation != g->graph_generation)
    return fail(FK_ERR_STATE, "fk_graph_replay: the grid was rebuilt or its tables "
                              "reallocated since the capture; capture again");
  FK_HIP(hipGraphLaunch(g->gexec, S(stream)));
  g->stat_set = kBatchedSet;
  g->stat_nq = g->batched_nq;
  return ok();
}

int64_t fk_idw_interp_workspace_bytes(int64_t nq, int64_t ns, int32_t k) {
  if (nq < 0 || ns < 0 || k < 1 || k > 64) return -1;
  return interp_workspace_bytes(nq, ns, k);
}

int fk_idw_interp_forward(const float *q, int64_t nq, const float *p, int64_t ns,
                          const int64_t *idx, int32_t k, const float *feat,
                          int32_t channels, float eps, float *w, float *out,
                          void *stream) {
  if (nq < 0 || ns < 0 || k < 1 || k > 64 || channels < 1 || nq * (int64_t)k > INT32_MAX ||
      ns > INT32_MAX || (nq > 0 && (!q || !idx || !out)) || (ns > 0 && (!p || !feat)) ||
      !std::isfinite(eps) || eps < 0.f)
    return fail(FK_ERR_INVALID_ARG, "fk_idw_interp_forward: bad arguments");
  FK_HIP(launch_interp_forward(q, nq, p, ns, idx, k, feat, channels, eps, w, out,
                               S(stream)));
  return ok();
}

int fk_idw_interp_backward(const float *q, int64_t nq, const float *p, int64_t ns,
                           const int64_t *idx, int32_t k, const float *feat,
                           int32_t channels, float eps, const float *grad_out,
                           float *grad_feat, float *grad_q, float *grad_p,
                           void *workspace, void *stream) {
  if (nq < 0 || ns < 0 || k < 1 || k > 64 || channels < 1 || nq * (int64_t)k > INT32_MAX ||
      ns > INT32_MAX || (nq > 0 && (!q || !idx || !grad_out)) ||
      (ns > 0 && (!p || !feat)) || !workspace || !std::isfinite(eps) || eps < 0.f)
    return fail(FK_ERR_INVALID_ARG, "fk_idw_interp_backward: bad arguments");
  FK_HIP(launch_interp_backward(q, nq, p, ns, idx, k, feat, channels, eps, grad_out,
                                grad_feat, grad_q, grad_p, workspace, S(stream)));
  return ok();
}

int64_t fk_voxel_workspace_bytes(int64_t n) {
  if (n < 0 || n > INT32_MAX / 2) return -1;
  return voxel_workspace_bytes(n);
}

int64_t fk_adaptive_workspace_bytes(int64_t n) {
  if (n < 1 || n > INT32_MAX / 2) return -1;
  return adaptive_workspace_bytes(n);
}

static_assert(FK_ADAPTIVE_STATS == kAdStats, "adaptive stats block");

int fk_adaptive_select(const int64_t *grid_coord, const int64_t *batch_idx,
                       const float *xyz, int64_t n, int32_t k, int32_t min_factor,
                       int32_t max_factor, int32_t max_shift, void *workspace,
                       int32_t *out_level, int64_t *stats, void *stream) {
  if (k < 1 || k > kMaxLaneK)
    return fail(FK_ERR_K_RANGE, "fk_adaptive_select: k = %d outside [1, %d]", k, kMaxLaneK);
  if (n < 1 || n > INT32_MAX / 2 || !grid_coord || !workspace || !stats || min_factor < 1 ||
      max_factor < min_factor || (int64_t)max_factor * k > INT32_MAX || max_shift < -1 ||
      max_shift > 16)
    return fail(FK_ERR_INVALID_ARG, "fk_adaptive_select: bad arguments");
  FK_HIP(launch_adaptive_select(grid_coord, batch_idx, xyz, n, min_factor * k,
                                max_factor * k, max_shift, workspace, out_level, stats,
                                S(stream)));
  return ok();
}

int fk_adaptive_query(void *workspace, int64_t n, int32_t k, float cut_r2,
                      const int64_t *stats, int64_t *out_idx, float *out_dist,
                      void *stream) {
  if (k < 1 || k > kMaxLaneK)
    return fail(FK_ERR_K_RANGE, "fk_adaptive_query: k = %d outside [1, %d]", k, kMaxLaneK);
  if (n < 1 || n > INT32_MAX / 2 || !workspace || !stats || !out_idx || std::isnan(cut_r2))
    return fail(FK_ERR_INVALID_ARG, "fk_adaptive_query: bad arguments");
  AdaptiveArgs a;
  memset(&a, 0, sizeof(a));
  FK_HIP(adaptive_tables(workspace, n, stats, &a));
  a.k = k;
  a.cut_r2 = cut_r2 > 0.f ? cut_r2 : 0.f;
  a.out_idx = out_idx;
  a.out_d = out_dist;
  hipStream_t s = S(stream);
  switch (kclass_of(k)) {
    case 8: FK_HIP(launch_adaptive_query<8>(a, s)); break;
    case 16: FK_HIP(launch_adaptive_query<16>(a, s)); break;
    case 32: FK_HIP(launch_adaptive_query<32>(a, s)); break;
    default: FK_HIP(launch_adaptive_query<64>(a, s)); break;
  }
  return ok();
}

int fk_voxel_first(const float *xyz, const int32_t *offsets, int32_t num_clouds,
                   int64_t n, float voxel_size, void *workspace,
                   int64_t *out_idx, int64_t *out_count, void *stream) {
  if (n < 0 || n > INT32_MAX / 2 || num_clouds < 1 || num_clouds > kMaxVoxelClouds ||
      (!offsets && num_clouds != 1) || (n > 0 && (!xyz || !out_idx)) ||
      !workspace || !out_count)
    return fail(FK_ERR_INVALID_ARG, "fk_voxel_first: bad arguments (num_clouds in "
                                    "[1, 32767], the reference's batch limit)");
  if (!(voxel_size > 0.f) || !std::isfinite(voxel_size))
    return fail(FK_ERR_INVALID_ARG,
                "fk_voxel_first: voxel_size must be finite and > 0");
  FK_HIP(launch_voxel_first(xyz, offsets, num_clouds, n, voxel_size, workspace,
                            out_idx, out_count, S(stream)));
  return ok();
}

int fk_voxel_first_noalloc(const float *xyz, const int32_t *offsets,
                           int32_t num_clouds, int64_t n, float voxel_size,
                           int64_t *out_idx, int64_t *out_count, void *stream) {
  if (n < 0 || num_clouds < 1 || num_clouds > kMaxVoxelClouds ||
      (!offsets && num_clouds != 1) || (n > 0 && (!xyz || !out_idx)) || !out_count)
    return fail(FK_ERR_INVALID_ARG, "fk_voxel_first_noalloc: bad arguments (num_clouds "
                                    "in [1, 32767])");
  if (n > FK_VOXEL_NOALLOC_MAX_POINTS)
    return fail(FK_ERR_CAPACITY, "fk_voxel_first_noalloc: %lld points is above the "
                                 "static workspace (%d points); use fk_voxel_first "
                                 "with a workspace", (long long)n,
                FK_VOXEL_NOALLOC_MAX_POINTS);
  if (!(voxel_size > 0.f) || !std::isfinite(voxel_size))
    return fail(FK_ERR_INVALID_ARG,
                "fk_voxel_first_noalloc: voxel_size must be finite and > 0");
  FK_HIP(launch_voxel_first_noalloc(xyz, offsets, num_clouds, n, voxel_size, out_idx,
                                    out_count, S(stream)));
  return ok();
}

int64_t fk_backward_workspace_bytes(int64_t nq, int64_t ns, int32_t k) {
  if (nq < 0 || ns < 0 || k < 1) return -1;
  return backward_workspace_bytes(nq, ns, k);
}

int fk_dist_backward(const float *q, int64_t nq, const float *p, int64_t ns,
                     const int64_t *idx, int32_t k, const float *grad_d,
                     float *grad_q, float *grad_p, void *workspace,
                     void *stream) {
  if (nq < 0 || ns < 0 || k < 1 || nq * (int64_t)k > INT32_MAX ||
      (nq > 0 && (!q || !idx || !grad_d)) || (ns > 0 && !p) ||
      (grad_p && ns > 0 && !workspace))
    return fail(FK_ERR_INVALID_ARG, "fk_dist_backward: bad arguments");
  FK_HIP(launch_dist_backward(q, nq, p, ns, idx, k, grad_d, grad_q, grad_p,
                              workspace, S(stream)));
  return ok();
}

int fk_idw_forward(const float *feat, int64_t ns, int32_t channels,
                   const int64_t *idx, const float *w, int64_t nq, int32_t k,
                   float *out, void *stream) {
  if (ns < 0 || nq < 0 || k < 1 || channels < 1 ||
      (nq > 0 && (!feat || !idx || !w || !out)))
    return fail(FK_ERR_INVALID_ARG, "fk_idw_forward: bad arguments");
  FK_HIP(launch_idw_forward(feat, ns, channels, idx, w, nq, k, out, S(stream)));
  return ok();
}

int fk_idw_backward(const float *feat, int64_t ns, int32_t channels,
                    const int64_t *idx, const float *w, int64_t nq, int32_t k,
                    const float *grad_out, float *grad_feat, float *grad_w,
                    void *workspace, void *stream) {
  if (ns < 0 || nq < 0 || k < 1 || channels < 1 || nq * (int64_t)k > INT32_MAX ||
      (nq > 0 && (!idx || !w || !grad_out)) || (grad_w && !feat) ||
      (grad_feat && ns > 0 && !workspace))
    return fail(FK_ERR_INVALID_ARG, "fk_idw_backward: bad arguments");
  FK_HIP(launch_idw_backward(feat, ns, channels, idx, w, nq, k, grad_out,
                             grad_feat, grad_w, workspace, S(stream)));
  return ok();
}

}  // extern "C"
