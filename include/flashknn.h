/*
 * flashknn.h - C ABI of the MI355X (gfx950) FlashKNN grid-kNN library.
 *
 * Plain pointers and sizes only (no torch types).  Every pointer named
 * "device" is HIP device memory; every entry point that takes a `stream`
 * enqueues its work on that HIP stream (hipStream_t passed as void*, NULL =
 * the null stream) and returns without a device-wide synchronisation.
 * Entry points return an fk_status (0 = FK_OK); nothing throws across this
 * ABI.  fk_last_error() returns a message for the last failure on the
 * calling thread.
 *
 * A grid handle (fk_grid) owns the device tables of one build: cell-sorted
 * support records, per-cell start offsets of a dense per-cloud cell box
 * (or, for boxes beyond 2^30 cells, per-(x, y)-column offsets with the
 * points of a column sorted by z: FK_BUILD_COLUMNS), per-cloud layout, and
 * optionally an attached, cell-sorted query list.
 * One handle must not be used from two streams at once.
 *
 * Reference interfaces replaced (reference = warriorsniu/FlashKNN_ExpRepo):
 *   fk_morton_encode          FlashKNN_Morton_Encode
 *                             FlashKNN/csrc/flash_knn_construction.cu:65-103
 *   fk_morton_neighbor_keys   FlashKNN_Morton_Neighbor_Keys
 *                             FlashKNN/csrc/flash_knn_construction.cu:105-126
 *   fk_grid_build_cells       the torch construction in FlashKNN.query
 *                             (argsort, unique_consecutive, searchsorted,
 *                             cumsum) FlashKNN/functions/FlashKnnWrapper.py:226-337
 *   fk_grid_build_xyz         the cell layout for exact queries of
 *                             pointops.knn_query (no reference construction:
 *                             the reference scans whole segments,
 *                             Pointcept/libs/pointops/src/knn_query/
 *                             knn_query_cuda_kernel.cu:60-104)
 *   fk_grid_attach_queries    query-side construction of selected_query /
 *                             back_query (FlashKnnWrapper.py:460-537,686-716)
 *   fk_query                  FlashKNN_Query_Dynamic_Load (api.cpp:9,
 *                             flash_knn_query_dynamic_load.cu:781-795),
 *                             FlashKNN_Selected_Query_DL (:1523-1641),
 *                             FlashKNN_Back_Query_DL / Nearest_Back_Query_DL
 *                             (:1373-1521) and pointops knn_query_cuda
 *                             (Pointcept/libs/pointops/src/knn_query/
 *                             knn_query_cuda.cpp:7)
 *   fk_batched_*              no reference counterpart (one persistent launch
 *                             over many clouds with per-cloud K, hipGraph
 *                             replay)
 *   fk_adaptive_select/query  AdaptiveNeighborhoodFlashKNN.query
 *                             (FlashKNN/functions/adaptive_octree.py:491-612:
 *                             build_adaptive_octree, select_adaptive_levels,
 *                             the flattened query :369-488)
 *   fk_dist_backward          no reference counterpart (FlashKNN.query is
 *                             @torch.no_grad, FlashKnnWrapper.py:165)
 *   fk_idw_forward/backward   pointops interpolation_forward/backward_cuda
 *                             (Pointcept/libs/pointops/src/interpolation/
 *                             interpolation_cuda_kernel.cu:5-33), with a
 *                             deterministic segmented backward
 */
#ifndef FLASHKNN_H
#define FLASHKNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FK_ABI_VERSION 1

typedef enum fk_status {
  FK_OK = 0,
  FK_ERR_INVALID_ARG = 1,    /* null pointer, negative size, bad enum     */
  FK_ERR_K_RANGE = 2,        /* k outside [1, FK_MAX_K]                   */
  FK_ERR_GRID_TOO_LARGE = 3, /* dense cell box above the cell capacity    */
  FK_ERR_DOMAIN = 4,         /* batch id / cell coordinate out of range   */
  FK_ERR_HIP = 5,            /* a HIP runtime call failed                 */
  FK_ERR_CAPACITY = 6,       /* FK_NO_SYNC build above reserved capacity  */
  FK_ERR_STATE = 7           /* handle not built / no queries attached    */
} fk_status;

/* fk_query: k <= 64 keeps each query's list in one lane's registers;
 * 65..128 (pointops nsample <= 128,
 * Pointcept/libs/pointops/src/knn_query/knn_query_cuda_kernel.cu) runs one
 * wave per query.  The batched kernel takes per-cloud k <= 64. */
#define FK_MAX_K 128
#define FK_MAX_BATCHED_K 64

/* fk_query_params.mode */
#define FK_MODE_CELLS 0 /* candidates = support points of the query's cloud
                           whose cell is within `radius_cells` of the
                           query's cell on every axis (1 => 27 cells): the
                           FlashKNN.query candidate set                    */
#define FK_MODE_EXACT 1 /* exact kNN within the query's cloud: cell scan,
                           per-query certification, shell expansion of the
                           uncertified (grids built by fk_grid_build_xyz)  */
/* fk_query_params.pad_mode: index written to slots that have no neighbour */
#define FK_PAD_SELF 0      /* the query's own id (FlashKNN)                */
#define FK_PAD_MINUS1 1    /* -1 (pointops, back_query)                   */
#define FK_PAD_PER_QUERY 2 /* per-query value given to attach_queries      */
/* fk_query_params.dist_mode */
#define FK_DIST_NONE 0 /* out_dist may be NULL                            */
#define FK_DIST_SQ 1   /* squared fp32 distance                           */
#define FK_DIST_SQRT 2 /* sqrt of the squared distance (pointops)         */
/* fk_query_params.query_set */
#define FK_QUERIES_SELF 0     /* every support point queries the grid      */
#define FK_QUERIES_ATTACHED 1 /* the list given to fk_grid_attach_queries   */
/* build flags */
#define FK_BUILD_NO_SYNC 1 /* never read sizes back to the host: use the
                              reserved capacity; overflow is reported by
                              fk_grid_status() (graph-capture safe).  Until
                              the next build, the query kernels of every
                              entry point (batched included) leave their
                              output arrays unwritten on a build that raised
                              an error                                      */
#define FK_BUILD_COLUMNS 2 /* column layout even when the dense cell box
                              fits (synchronising builds only).  A
                              synchronising build whose dense boxes exceed
                              2^30 cells takes this layout by itself: one
                              table entry per (x, y) column, points sorted by
                              z inside it; results are the same, lookups
                              search the column                            */

typedef struct fk_query_params {
  int32_t k;            /* neighbours per query, 1..FK_MAX_K               */
  int32_t mode;         /* FK_MODE_*                                        */
  int32_t radius_cells; /* FK_MODE_CELLS: block half-width (>= 1)          */
  float cut_r2;         /* keep d < cut_r2 (strict); INFINITY = no cut      */
  int32_t idx_64;       /* 1: out_idx is int64, 0: int32                    */
  int32_t pad_mode;     /* FK_PAD_*                                         */
  int32_t dist_mode;    /* FK_DIST_*                                        */
  float pad_dist;       /* value written to out_dist for padded slots      */
  int32_t query_set;    /* FK_QUERIES_*                                     */
  int32_t passes;       /* 0 = all; else a mask of the launches fk_query
                           enqueues: 1 counter reset, 2 first (cell block)
                           pass, 4 expansion pass (FK_MODE_EXACT).  Lets a
                           profiler time one kernel alone.  With 0 the
                           counters are reset only when needed (each first
                           pass clears the next call's set); a caller that
                           captures fk_query into a graph passes 7 so that
                           the reset is part of every replay.               */
  int32_t flags;        /* FK_QUERY_* hints (results never change)          */
  int32_t reserved;
} fk_query_params;

/* fk_query_params.flags */
#define FK_QUERY_PRUNE 1 /* after the query's own column, skip neighbour
                            cells whose nearest possible point is beyond the
                            current k-th distance (grids built from xyz, k
                            <= 16); faster when the 27-cell block holds many
                            more candidates than k                         */
#define FK_QUERY_ONE_LANE 2 /* first pass with one lane (64 queries) per
                               wave instead of two lanes per query, for
                               k <= 32 (fk_query; the batched kernel keeps
                               two).  Without it, a launch too small to
                               fill the GPU (under 2,048 waves) widens
                               its groups to 4, 8 or 16 lanes per query;
                               results never depend on the group size     */
#define FK_QUERY_UNIFORM 4 /* first pass in uniform tiles.  Without it, a
                              self-query launch of 65,536 queries or more
                              that fits one resident round runs a plan made
                              once per build: super-tiles of 64 queries
                              whose 27-cell candidate counts are high get up
                              to 8 lanes per query, light ones down to 1, so
                              that no wave's scan runs far longer than the
                              rest (results never change)                 */
#define FK_QUERY_BATCH_SMALL 8 /* candidate batch of the first pass: the
                                  records each lane loads per scan round
                                  before the wave checks its LDS key
                                  buffers.  Default 8 (k <= 16) / 4 (k in
                                  17..32); SMALL halves it, LARGE takes
                                  12 / 8 (more loads in flight, fewer
                                  checks, less buffer room: more merges).
                                  First passes of k <= 32 on dense grids
                                  (fk_query; the batched kernels keep the
                                  default); not both flags.  Compiled variants chosen at
                                  run time, like the reference's tile
                                  variants (FlashKNN/csrc/flash_knn_query_
                                  dynamic_load.cu:903-945); the autotuner
                                  picks one per bucket.  Results never
                                  change                                  */
#define FK_QUERY_BATCH_LARGE 16

typedef struct fk_grid fk_grid;

/* ------------------------------------------------------------- library */
const char *fk_status_string(int status);
const char *fk_last_error(void);
int fk_abi_version(void);

/* ------------------------------------------------ Morton keys (drop-in) */
/* keys[i] = 48-bit Morton key of grid_coord[i] (16 bits per axis, x at bit
 * 3i+2, y at 3i+1, z at 3i) | batch_idx[i] << 48.  All pointers device. */
int fk_morton_encode(const int64_t *grid_coord, const int64_t *batch_idx,
                     int64_t n, int64_t *keys, void *stream);
/* keys[27p+j] = key of grid_down[p] + offset j, offsets [0,-1,1]^3 in ij
 * order (centre first). */
int fk_morton_neighbor_keys(const int64_t *grid_down,
                            const int64_t *parent_batch, int64_t num_parents,
                            int64_t *keys, void *stream);

/* ---------------------------------------------------------------- grids */
int fk_grid_create(fk_grid **out);
int fk_grid_destroy(fk_grid *grid);
/* Pre-size the device tables (needed before FK_BUILD_NO_SYNC builds).  A
 * reserve that grows the build's tables leaves the handle unbuilt (their
 * contents are gone), one that grows the query tables leaves no queries
 * attached: build (and attach) again before querying. */
int fk_grid_reserve(fk_grid *grid, int64_t max_points, int64_t max_cells,
                    int32_t max_clouds, int64_t max_queries);

/* Build from integer voxel coordinates (FlashKNN.query layout):
 * cell = grid_coord >> num_down (int64[n,3], device), cloud = batch_idx[i]
 * (int64[n], device, NULL => one cloud) in [0, num_clouds).  xyz (float[n,3],
 * device) gives the coordinates distances are taken on; NULL => grid_coord
 * converted to float.  Support ids are 0..n-1. */
int fk_grid_build_cells(fk_grid *grid, const float *xyz,
                        const int64_t *grid_coord, const int64_t *batch_idx,
                        int64_t n, int32_t num_clouds, int32_t num_down,
                        int32_t flags, void *stream);

/* A build that returns an error leaves the handle unbuilt: until a build
 * succeeds, queries, attaches and table reads return FK_ERR_STATE (the
 * failed build may already have reallocated tables that the earlier build
 * described).  An attach that returns an error leaves no queries attached.
 * (A no-sync build's device-side errors surface later, see
 * FK_BUILD_NO_SYNC.) */

/* Build from coordinates (exact-query layout): point i belongs to cloud c
 * with offsets[c-1] <= i < offsets[c] (int32 cumulative ends, device);
 * cell = floor(xyz * inv_cell[c]) (float[num_clouds], device).  Support ids
 * are 0..n-1. */
int fk_grid_build_xyz(fk_grid *grid, const float *xyz, const int32_t *offsets,
                      int32_t num_clouds, int64_t n, const float *inv_cell,
                      int32_t flags, void *stream);

/* Cell side per cloud for exact queries, from the cloud's bounding-box
 * density: side = scale * (k * volume / n)^(1/3) (degenerate axes count with
 * 1% of the largest extent; at most ~64 cells per point).  k from
 * k_per_cloud (int32[num_clouds], device) or the scalar k.  Writes
 * side_out (nullable) and inv_out = 1/side (float[num_clouds], device); uses
 * the handle's per-cloud scratch; no host synchronisation.  No reference
 * counterpart (the reference's caller picks voxel size and num_down). */
int fk_grid_auto_cell(fk_grid *grid, const float *xyz, const int32_t *offsets,
                      int32_t num_clouds, int64_t n, const int32_t *k_per_cloud,
                      int32_t k, float scale, float *side_out, float *inv_out,
                      void *stream);
/* Density feedback on a grid built by fk_grid_build_xyz: per cloud, the mean
 * 27-cell candidate count m of the self queries gives
 * side *= clamp((ratio * k / m)^(1/exponent), 1/4, 4) and inv = 1/side, in
 * place (device), with ratio = target_per_cloud[c] (float[num_clouds],
 * device, nullable) or target_ratio; rebuild with the new inv afterwards. */
int fk_grid_adapt_cell(fk_grid *grid, const int32_t *k_per_cloud, int32_t k,
                       float target_ratio, const float *target_per_cloud,
                       float exponent, float *side_io, float *inv_io,
                       void *stream);
/* Per-cloud sum of the 27-cell candidate counts of the self queries
 * (int64[num_clouds], device; no host synchronisation). */
int fk_grid_cloud_candidates(fk_grid *grid, int64_t *out, void *stream);

/* Attach a query list in the grid's cell space.  For grids built from
 * cells: qcell (int64[nq,3], device) are voxel coordinates (shifted by the
 * grid's num_down) and qbatch (int64[nq], device, nullable) the cloud.  For
 * grids built from xyz: qcell must be NULL; the cell is floor(qxyz*inv) and
 * the cloud comes from qoffsets (int32 cumulative ends over the queries).
 * Results of query i are written to output row i.  qpad (int32[nq], device,
 * nullable) supplies the FK_PAD_PER_QUERY value. */
int fk_grid_attach_queries(fk_grid *grid, const float *qxyz,
                           const int64_t *qcell, const int64_t *qbatch,
                           const int32_t *qoffsets, int64_t nq,
                           const int32_t *qpad, int32_t flags, void *stream);

/* Device status word of the last build (FK_OK or the first error a kernel
 * raised); reading it synchronises the stream and makes the cell count of
 * an FK_BUILD_NO_SYNC build known to fk_grid_info. */
int fk_grid_status(fk_grid *grid, void *stream);
/* Host-side sizes of the last build (points, clouds, cells; -1 if unknown
 * in FK_BUILD_NO_SYNC mode). */
int fk_grid_info(fk_grid *grid, int64_t *num_points, int32_t *num_clouds,
                 int64_t *num_cells, int64_t *num_queries);
/* Device pointers of the build (for tests and the backward): sorted support
 * records float4 (x,y,z,id bits), cell start table int32[num_cells+1] (one
 * entry per cell; per (x, y) column in the column layout, FK_BUILD_COLUMNS,
 * where fk_grid_info's num_cells counts columns). */
int fk_grid_tables(fk_grid *grid, const void **sorted_records,
                   const int32_t **cell_start);

/* ---------------------------------------------------------------- query */
/* Row r of out_idx/out_dist (k entries, ascending by (d, id)) belongs to
 * support point r (FK_QUERIES_SELF) or attached query r.  out_idx is
 * int32/int64 per params->idx_64; out_dist may be NULL when dist_mode is
 * FK_DIST_NONE. */
int fk_query(fk_grid *grid, const fk_query_params *params, void *out_idx,
             float *out_dist, void *stream);
/* Statistics of the last FK_MODE_EXACT query: queries certified by the
 * first pass (the query count minus those left pending), and those
 * resolved by shell expansion; 0 and 0 after a 27-cell query.
 * Synchronises. */
int fk_query_stats(fk_grid *grid, int64_t *first_pass_certified,
                   int64_t *expanded, void *stream);
/* Sum over the query set of the number of candidates in the block of
 * half-width radius_cells (FlashKNN.query's traverse_info statistic,
 * FlashKnnWrapper.py:351-352, is this sum / N).  Synchronises. */
int fk_candidate_count(fk_grid *grid, int32_t radius_cells, int32_t query_set,
                       int64_t *total, void *stream);
/* Per-query candidate count |C(q)| (int32[nq], device, row order of the
 * query set); the node statistic the reference's adaptive level selection
 * thresholds (FlashKNN/functions/adaptive_octree.py:204-208,312-351). */
int fk_point_candidates(fk_grid *grid, int32_t radius_cells, int32_t query_set,
                        int32_t *out, void *stream);

/* ------------------------------------------- batched persistent query */
/* All clouds of a grid built by fk_grid_build_xyz or fk_grid_build_cells,
 * with a per-cloud k, in one call: one tile-count launch, then per K class
 * up to params->k's (8 / 16 / 32 / 64) one persistent kernel with its own
 * register budget and tile queue followed by the expansion of that class's
 * uncertified queries; the classes are forked onto the handle's auxiliary
 * streams and joined back onto `stream`.
 * k_per_cloud: int32[num_clouds] (device).  out_offsets: int64[num_clouds]
 * (device) start of cloud c's rows in out_idx/out_dist, in elements
 * (row r of cloud c lives at out_offsets[c] + (r - first point of c) * k_c).
 * params->k must be the maximum of k_per_cloud. */
int fk_batched_query(fk_grid *grid, const fk_query_params *params,
                     const int32_t *k_per_cloud, const int64_t *out_offsets,
                     void *out_idx, float *out_dist, void *stream);
/* Capture fk_batched_query into a hipGraph owned by the handle (the
 * arguments are frozen); fk_graph_replay launches it on `stream`.  The
 * graph holds the handle's table pointers and the sizes of the build it
 * was captured on: after any later build on the handle, or any call that
 * reallocated its tables (a larger fk_grid_reserve, fk_query or
 * fk_batched_query needing more pending-query room), fk_graph_replay
 * launches nothing and returns FK_ERR_STATE; capture again. */
int fk_batched_capture(fk_grid *grid, const fk_query_params *params,
                       const int32_t *k_per_cloud, const int64_t *out_offsets,
                       void *out_idx, float *out_dist, void *stream);
int fk_graph_replay(fk_grid *grid, void *stream);

/* ------------------------------------------------------------- backward */
/* Gradient of d_ij = ||q_i - p_idx(i,j)||^2 given grad_d [nq,k]:
 * grad_q[i] = sum_j 2 g_ij (q_i - p_j)  (float[nq,3], overwritten)
 * grad_p[m] = -sum_{idx=m} 2 g_ij (q_i - p_m) (float[ns,3], overwritten),
 * summed in (i, j) order per m (deterministic segmented reduction over a
 * reverse CSR).  idx is int64 [nq,k]; entries < 0 are skipped.  workspace
 * must hold fk_backward_workspace_bytes(nq, ns, k) bytes (device). */
int64_t fk_backward_workspace_bytes(int64_t nq, int64_t ns, int32_t k);
int fk_dist_backward(const float *q, int64_t nq, const float *p, int64_t ns,
                     const int64_t *idx, int32_t k, const float *grad_d,
                     float *grad_q, float *grad_p, void *workspace,
                     void *stream);

/* IDW: out[i,:] = sum_j w[i,j] * feat[idx[i,j],:]  (w given, [nq,k]). */
int fk_idw_forward(const float *feat, int64_t ns, int32_t channels,
                   const int64_t *idx, const float *w, int64_t nq, int32_t k,
                   float *out, void *stream);
/* grad_feat[m,:] = sum_{idx=m} w_ij grad_out[i,:] (segmented, deterministic);
 * grad_w[i,j] = <grad_out[i,:], feat[idx_ij,:]>.  Either output may be NULL.
 * workspace as for fk_dist_backward. */
int fk_idw_backward(const float *feat, int64_t ns, int32_t channels,
                    const int64_t *idx, const float *w, int64_t nq, int32_t k,
                    const float *grad_out, float *grad_feat, float *grad_w,
                    void *workspace, void *stream);

/* Fused IDW interpolation over neighbour lists (no counterpart in the
 * reference, whose pointops interpolation differentiates features only,
 * Pointcept/libs/pointops/functions/interpolation.py:8-56):
 *   d_ij = ||p_idx(i,j) - q_i||^2 (the query kernels' fp32 form),
 *   r_ij = 1 / (d_ij + eps) (0 for idx < 0), w_ij = r_ij / sum_l r_il,
 *   out[i,:] = sum_j w_ij feat[idx_ij,:]    (out float[nq,C]).
 * k <= 64.  w (float[nq,k]) may be NULL.  The backward writes any non-NULL
 * of grad_feat [ns,C], grad_q [nq,3], grad_p [ns,3] (overwritten; the
 * support-side sums run over a reverse CSR in entry order: reproducible);
 * workspace: fk_idw_interp_workspace_bytes(nq, ns, k) bytes (device). */
int64_t fk_idw_interp_workspace_bytes(int64_t nq, int64_t ns, int32_t k);
int fk_idw_interp_forward(const float *q, int64_t nq, const float *p, int64_t ns,
                          const int64_t *idx, int32_t k, const float *feat,
                          int32_t channels, float eps, float *w, float *out,
                          void *stream);
int fk_idw_interp_backward(const float *q, int64_t nq, const float *p, int64_t ns,
                           const int64_t *idx, int32_t k, const float *feat,
                           int32_t channels, float eps, const float *grad_out,
                           float *grad_feat, float *grad_q, float *grad_p,
                           void *workspace, void *stream);

/* ------------------------------------------- adaptive neighbourhoods */
/* AdaptiveNeighborhoodFlashKNN.query (FlashKNN/functions/adaptive_octree.py
 * :491-612) for k <= 64, in two calls with no host synchronisation:
 *
 * fk_adaptive_select: shifts each batch's int64 cell coordinates to start
 * at 0 (:129-146; spans must stay below 2^16, batch indices in [0, 2^15)),
 * takes the coarsest level max_shift (-1: the bit length of the largest
 * shifted coordinate) and, per point, the reference's coarse-to-fine rule
 * over the 27-node candidate counts of levels max_shift .. 0 (:312-351:
 * descend while the count exceeds max_factor * k, accept a finer level
 * only while its count is at least min_factor * k).  grid_coord int64[n,3],
 * batch_idx int64[n] or NULL (one batch), xyz float[n,3] or NULL (the
 * shifted cell coordinates are the positions), all device.  Writes the
 * selected level per point to out_level (int32[n], device, nullable) and
 * the statistics block stats (int64[FK_ADAPTIVE_STATS], device): [0] a
 * device-side status (FK_ERR_DOMAIN), [1] its reason (1 batch index, 2
 * span), [2] the largest shifted coordinate, [3] the coarsest level, then
 * per level s = 0..16 at [4 + 4s]: points that selected s and the sum, min
 * and max of their counts.
 *
 * fk_adaptive_query: the k nearest (fp32 squared distance, ties to the
 * smaller index, d < cut_r2) among the points of the same batch in the 27
 * nodes around each point's node at its selected level (:369-488); rows in
 * input order, int64 out_idx[n,k], float out_dist[n,k] (nullable); slots
 * without a candidate hold (the point's own index, +inf).  Reads the tables
 * fk_adaptive_select left in the workspace; writes nothing when stats[0]
 * holds an error.
 *
 * workspace (device) must hold fk_adaptive_workspace_bytes(n) bytes.
 * k outside [1, 64] returns FK_ERR_K_RANGE. */
#define FK_ADAPTIVE_STATS 72
int64_t fk_adaptive_workspace_bytes(int64_t n);
int fk_adaptive_select(const int64_t *grid_coord, const int64_t *batch_idx,
                       const float *xyz, int64_t n, int32_t k,
                       int32_t min_factor, int32_t max_factor,
                       int32_t max_shift, void *workspace, int32_t *out_level,
                       int64_t *stats, void *stream);
int fk_adaptive_query(void *workspace, int64_t n, int32_t k, float cut_r2,
                      const int64_t *stats, int64_t *out_idx, float *out_dist,
                      void *stream);

/* -------------------------------------------- voxel-first subsampling */
/* Replaces _voxel_first_cuda (Networks/hierarchy.py:23-36), the step that
 * builds every coarser level of the hierarchy before the kNN path.  Voxel
 * v = floor(x / voxel_size) per axis (correctly rounded fp32 quotient, as
 * torch divides a device tensor by a Python scalar), keyed by its low 16
 * bits per axis (FlashKNN/functions/z_order.py:56-91) and by its cloud;
 * the voxel's representative is its smallest point index.  Writes the kept
 * indices ascending to out_idx (int64[n], device) and their count to
 * *out_count (int64, device; -status on a device-side error: a coordinate
 * outside |v| < 2^31 or not finite, a point past offsets[num_clouds-1]).
 * offsets: int32[num_clouds] cumulative ends (device) or NULL for one cloud;
 * num_clouds <= 32767 (the reference's batch limit; the key holds the cloud
 * in 15 bits), else FK_ERR_INVALID_ARG.
 * No host synchronisation; workspace (device) must hold
 * fk_voxel_workspace_bytes(n) bytes. */
int64_t fk_voxel_workspace_bytes(int64_t n);
int fk_voxel_first(const float *xyz, const int32_t *offsets, int32_t num_clouds,
                   int64_t n, float voxel_size, void *workspace,
                   int64_t *out_idx, int64_t *out_count, void *stream);

/* fk_voxel_first without a workspace argument and without any allocation:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The workspace is a static array of the library, one
 * per GPU (36 MiB, held until the process ends), sized for
 * n <= FK_VOXEL_NOALLOC_MAX_POINTS; larger n returns FK_ERR_CAPACITY and
 * launches nothing.  Results are those of fk_voxel_first.  Every call on
 * one GPU uses that GPU's one array: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialised by the caller. */
#define FK_VOXEL_NOALLOC_MAX_POINTS 1048576
int fk_voxel_first_noalloc(const float *xyz, const int32_t *offsets,
                           int32_t num_clouds, int64_t n, float voxel_size,
                           int64_t *out_idx, int64_t *out_count, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FLASHKNN_H */
