/*
 * CPU oracle for the FlashKNN grid-kNN path.  TEST INFRASTRUCTURE ONLY:
 * nothing in the shipped library links, loads or calls this file.  It is the
 * checker that the HIP kernels are compared against.
 *
 * Built with -O2 -ffp-contract=off -fno-fast-math so every fp32 expression is
 * evaluated as written (no fused multiply-add, no reassociation).
 *
 * Contents (each function cites the reference code it restates):
 *   fko_morton_encode / fko_morton_neighbor_keys
 *       FlashKNN/csrc/flash_knn_construction.cu:13-63 (48-bit key, 16 bits
 *       per axis, x at bit 3i+2, y at 3i+1, z at 3i, batch << 48; neighbour
 *       offsets ordered [0,-1,1]^3 with ij indexing).
 *   fko_grid_knn
 *       Candidate-set semantics of FlashKNN_Query_dynamic_load_kernel
 *       (FlashKNN/csrc/flash_knn_query_dynamic_load.cu:469-652) as driven by
 *       FlashKNN.query / selected_query / back_query
 *       (FlashKNN/functions/FlashKnnWrapper.py:165-849): the candidates of a
 *       query are the support points of the same batch whose search cell
 *       differs from the query's by at most R per axis (R = 1: 27 cells);
 *       squared distance fl(fl(fl(dx*dx)+fl(dy*dy))+fl(dz*dz)) with
 *       dx = fl(p.x - q.x) (dynamic_load.cu:601-604); a candidate enters only
 *       if d < cut_r2 (strict, dynamic_load.cu:606); the K smallest by
 *       (d, index) are returned ascending; missing slots are (pad, +inf)
 *       (init at dynamic_load.cu:562).
 *   fko_brute_knn
 *       Exact per-segment kNN of pointops.knn_query
 *       (Pointcept/libs/pointops/src/knn_query/knn_query_cuda_kernel.cu:60-104,
 *       offsets = cumulative segment ends, :70-76), same distance expression
 *       ((a+b)+c, :92), (d, index) order, padding (-1, +inf); optional fp64
 *       ranking (the FP64 oracle).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- Morton */

static int64_t morton_key1(int64_t x, int64_t y, int64_t z, int64_t b) {
  uint32_t ux = (uint32_t)x & 0xffffu, uy = (uint32_t)y & 0xffffu,
           uz = (uint32_t)z & 0xffffu;
  uint64_t key = 0;
  for (int bit = 0; bit < 16; ++bit) {
    key |= (uint64_t)((ux >> bit) & 1u) << (3 * bit + 2);
    key |= (uint64_t)((uy >> bit) & 1u) << (3 * bit + 1);
    key |= (uint64_t)((uz >> bit) & 1u) << (3 * bit);
  }
  key |= (uint64_t)b << 48;
  return (int64_t)key;
}

void fko_morton_encode(const int64_t *grid, const int64_t *batch, int64_t n,
                       int64_t *out) {
  for (int64_t i = 0; i < n; ++i)
    out[i] = morton_key1(grid[3 * i], grid[3 * i + 1], grid[3 * i + 2],
                         batch ? batch[i] : 0);
}

static int nbr_off(int pos) { return pos == 0 ? 0 : (pos == 1 ? -1 : 1); }

void fko_morton_neighbor_keys(const int64_t *grid_down, const int64_t *pbatch,
                              int64_t P, int64_t *out) {
  for (int64_t p = 0; p < P; ++p)
    for (int nb = 0; nb < 27; ++nb)
      out[p * 27 + nb] = morton_key1(grid_down[3 * p] + nbr_off(nb / 9),
                                     grid_down[3 * p + 1] + nbr_off((nb / 3) % 3),
                                     grid_down[3 * p + 2] + nbr_off(nb % 3),
                                     pbatch ? pbatch[p] : 0);
}

/* ------------------------------------------------------------- distances */

static inline float d32(const float *p, const float *q) {
  float dx = p[0] - q[0];
  float dy = p[1] - q[1];
  float dz = p[2] - q[2];
  float xx = dx * dx;
  float yy = dy * dy;
  float zz = dz * dz;
  float s = xx + yy;
  return s + zz;
}

static inline double d64(const float *p, const float *q) {
  double dx = (double)p[0] - (double)q[0];
  double dy = (double)p[1] - (double)q[1];
  double dz = (double)p[2] - (double)q[2];
  return dx * dx + dy * dy + dz * dz;
}

static inline uint64_t make_key(float d, uint32_t idx) {
  uint32_t b;
  memcpy(&b, &d, 4);
  return ((uint64_t)b << 32) | idx;
}

static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

typedef struct {
  double d;
  int64_t i;
} dkey;

static int cmp_dkey(const void *a, const void *b) {
  const dkey *x = (const dkey *)a, *y = (const dkey *)b;
  if (x->d < y->d) return -1;
  if (x->d > y->d) return 1;
  return x->i < y->i ? -1 : (x->i > y->i ? 1 : 0);
}

/* --------------------------------------------------- grid candidate kNN */

typedef struct {
  int64_t b, x, y, z;
} cellkey;

static const cellkey *g_sort_keys; /* qsort context (single-threaded sort) */

static int cmp_cellkey(const cellkey *a, const cellkey *b) {
  if (a->b != b->b) return a->b < b->b ? -1 : 1;
  if (a->x != b->x) return a->x < b->x ? -1 : 1;
  if (a->y != b->y) return a->y < b->y ? -1 : 1;
  if (a->z != b->z) return a->z < b->z ? -1 : 1;
  return 0;
}

static int cmp_order(const void *a, const void *b) {
  int64_t i = *(const int64_t *)a, j = *(const int64_t *)b;
  int c = cmp_cellkey(&g_sort_keys[i], &g_sort_keys[j]);
  if (c) return c;
  return i < j ? -1 : (i > j ? 1 : 0);
}

/* first position in order[] whose key >= k */
static int64_t lower_bound(const int64_t *order, const cellkey *keys, int64_t n,
                           const cellkey *k) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (cmp_cellkey(&keys[order[mid]], k) < 0)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

/*
 * Support S: ns points (sxyz f32[ns*3], scell i64[ns*3], sbatch i64[ns] or
 * NULL).  Queries Q: nq points (qxyz, qcell, qbatch or NULL, qpad i64[nq]).
 * Output idx i64[nq*K] (support indices), d f32[nq*K].  Returns 0.
 */
int fko_grid_knn(int64_t ns, const float *sxyz, const int64_t *scell,
                 const int64_t *sbatch, int64_t nq, const float *qxyz,
                 const int64_t *qcell, const int64_t *qbatch,
                 const int64_t *qpad, int R, int K, float cut_r2,
                 int64_t *out_idx, float *out_d) {
  cellkey *keys = (cellkey *)malloc(sizeof(cellkey) * (size_t)(ns ? ns : 1));
  int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ns ? ns : 1));
  for (int64_t i = 0; i < ns; ++i) {
    keys[i].b = sbatch ? sbatch[i] : 0;
    keys[i].x = scell[3 * i];
    keys[i].y = scell[3 * i + 1];
    keys[i].z = scell[3 * i + 2];
    order[i] = i;
  }
  g_sort_keys = keys;
  qsort(order, (size_t)ns, sizeof(int64_t), cmp_order);
  const uint64_t cutkey = make_key(cut_r2, 0u);
  const uint32_t inf_bits = 0x7f800000u;

#pragma omp parallel
  {
    size_t cap = 1024;
    uint64_t *cand = (uint64_t *)malloc(sizeof(uint64_t) * cap);
#pragma omp for schedule(dynamic, 64)
    for (int64_t q = 0; q < nq; ++q) {
      const float *qp = qxyz + 3 * q;
      int64_t b = qbatch ? qbatch[q] : 0;
      int64_t cx = qcell[3 * q], cy = qcell[3 * q + 1], cz = qcell[3 * q + 2];
      size_t n = 0;
      for (int64_t dx = -R; dx <= R; ++dx)
        for (int64_t dy = -R; dy <= R; ++dy) {
          cellkey lo = {b, cx + dx, cy + dy, cz - R};
          cellkey hi = {b, cx + dx, cy + dy, cz + R + 1};
          int64_t a = lower_bound(order, keys, ns, &lo);
          int64_t e = lower_bound(order, keys, ns, &hi);
          for (int64_t t = a; t < e; ++t) {
            int64_t s = order[t];
            float d = d32(sxyz + 3 * s, qp);
            uint64_t k = make_key(d, (uint32_t)s);
            if (!(k < cutkey)) continue;
            if (n == cap) {
              cap *= 2;
              cand = (uint64_t *)realloc(cand, sizeof(uint64_t) * cap);
            }
            cand[n++] = k;
          }
        }
      qsort(cand, n, sizeof(uint64_t), cmp_u64);
      for (int j = 0; j < K; ++j) {
        if ((size_t)j < n) {
          uint32_t hi = (uint32_t)(cand[j] >> 32);
          float d;
          memcpy(&d, &hi, 4);
          out_idx[q * K + j] = (int64_t)(uint32_t)(cand[j] & 0xffffffffu);
          out_d[q * K + j] = d;
        } else {
          float inf;
          memcpy(&inf, &inf_bits, 4);
          out_idx[q * K + j] = qpad[q];
          out_d[q * K + j] = inf;
        }
      }
    }
    free(cand);
  }
  free(keys);
  free(order);
  return 0;
}

/* ------------------------------------------------------ exact brute force */

/*
 * Segments: support offsets soff[nb] and query offsets qoff[nb] are
 * cumulative END indices (pointops convention).  For each query, the K
 * smallest (d, support index) over its segment.  use_f64 = 0: fp32 d as
 * above, out_d is float*.  use_f64 = 1: ranking by the fp64 distance of the
 * fp32 inputs, out_d is double*.  Padding: (pad_idx, +inf).
 */
/* Insert key k into best[0..*n) (ascending, at most K entries), dropping
 * the largest when full. */
static inline void keep_smallest(uint64_t *best, size_t *n, int K, uint64_t k) {
  size_t m = *n;
  if (m == (size_t)K) {
    if (k >= best[K - 1]) return;
    m = (size_t)K - 1;
  } else {
    *n = m + 1;
  }
  while (m > 0 && best[m - 1] > k) {
    best[m] = best[m - 1];
    --m;
  }
  best[m] = k;
}

int fko_brute_knn(int64_t ns, const float *sxyz, const int64_t *soff,
                  int64_t nq, const float *qxyz, const int64_t *qoff, int nb,
                  int K, float cut_r2, int64_t pad_idx, int use_f64,
                  int64_t *out_idx, void *out_d) {
  const uint64_t cutkey = make_key(cut_r2, 0u);
  (void)ns;
  (void)nq;
  for (int s = 0; s < nb; ++s) {
    int64_t s0 = s ? soff[s - 1] : 0, s1 = soff[s];
    int64_t q0 = s ? qoff[s - 1] : 0, q1 = qoff[s];
    int64_t m = s1 - s0;
#pragma omp parallel
    {
      uint64_t *cand = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(K > 0 ? K : 1));
      dkey *dc = use_f64 ? (dkey *)malloc(sizeof(dkey) * (size_t)(m ? m : 1)) : NULL;
#pragma omp for schedule(dynamic, 64)
      for (int64_t q = q0; q < q1; ++q) {
        const float *qp = qxyz + 3 * q;
        size_t n = 0;
        if (!use_f64) {
          /* the K smallest keys, kept sorted (keys are distinct: the index
             is in the low bits), equal to the first K of a full sort */
          for (int64_t t = s0; t < s1; ++t) {
            uint64_t k = make_key(d32(sxyz + 3 * t, qp), (uint32_t)t);
            if (k < cutkey) keep_smallest(cand, &n, K, k);
          }
          float *od = (float *)out_d;
          for (int j = 0; j < K; ++j) {
            if ((size_t)j < n) {
              uint32_t hi = (uint32_t)(cand[j] >> 32);
              memcpy(&od[q * K + j], &hi, 4);
              out_idx[q * K + j] = (int64_t)(uint32_t)(cand[j] & 0xffffffffu);
            } else {
              od[q * K + j] = INFINITY;
              out_idx[q * K + j] = pad_idx;
            }
          }
        } else {
          for (int64_t t = s0; t < s1; ++t) {
            double d = d64(sxyz + 3 * t, qp);
            if (d < (double)cut_r2) {
              dc[n].d = d;
              dc[n].i = t;
              ++n;
            }
          }
          qsort(dc, n, sizeof(dkey), cmp_dkey);
          double *od = (double *)out_d;
          for (int j = 0; j < K; ++j) {
            if ((size_t)j < n) {
              od[q * K + j] = dc[j].d;
              out_idx[q * K + j] = dc[j].i;
            } else {
              od[q * K + j] = INFINITY;
              out_idx[q * K + j] = pad_idx;
            }
          }
        }
      }
      free(cand);
      free(dc);
    }
  }
  return 0;
}

int fko_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
