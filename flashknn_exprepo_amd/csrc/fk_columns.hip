// Column layout of a grid build: points ordered by (column, z).
//
// A dense box has one start-table entry per cell; a cloud whose box would
// need more than the table can hold (fine cells over a large extent, e.g.
// 2 cm voxels over a 100 m LiDAR sweep: ~5e9 cells) is laid out by (x, y)
// columns instead.  The counting sort by column gives the start table; a
// radix sort of (column << depth_bits | local z) keys then orders the points
// inside each column by z, so the query kernels find the run of a z segment
// by binary search in the column (fk_query.hip, column_range).  The
// candidate sets are those of the dense layout: only the lookup differs.
#include "fk_internal.h"

#include <rocprim/device/device_radix_sort.hpp>

namespace fk {

namespace {

constexpr int kBlock = 256;

inline unsigned blocks_of(int64_t n) {
  const int64_t b = (n + kBlock - 1) / kBlock;
  return (unsigned)(b < 1 ? 1 : b);
}

// key = table entry << depth_bits | local z; payload = input index
__global__ void k_column_keys(const int4 *pt_cell, const int32_t *pt_lin,
                              int64_t n, const CloudInfo *clouds,
                              int32_t depth_bits, uint64_t *keys,
                              int32_t *perm, const int32_t *scalars) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t key = 0;
  if (scalars[kStatus] == 0) {
    const int4 c = pt_cell[i];
    const uint32_t lz = (uint32_t)(c.z - clouds[c.w].minz);
    key = ((uint64_t)(uint32_t)pt_lin[i] << depth_bits) | lz;
  }
  keys[i] = key;
  perm[i] = (int32_t)i;
}

// Sorted position p takes input point perm[p].
__global__ void k_scatter_perm(const float *xyz, const int64_t *grid, int64_t n,
                               const int32_t *perm, const int4 *pt_cell,
                               const CloudInfo *clouds, float4 *srec,
                               int4 *scell, const int32_t *scalars) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n || scalars[kStatus] != 0) return;
  const int64_t i = perm[p];
  const int4 c = pt_cell[i];
  const CloudInfo ci = clouds[c.w];
  float x, y, z;
  if (xyz) {
    x = xyz[3 * i]; y = xyz[3 * i + 1]; z = xyz[3 * i + 2];
  } else {
    x = (float)grid[3 * i]; y = (float)grid[3 * i + 1]; z = (float)grid[3 * i + 2];
  }
  srec[p] = make_float4(x, y, z, __int_as_float((int)i));
  scell[p] = make_int4(c.x - ci.minx, c.y - ci.miny, c.z - ci.minz, c.w);
}

}  // namespace

hipError_t column_sort_tmp_bytes(int64_t n, int key_bits, size_t *bytes) {
  *bytes = 0;
  return rocprim::radix_sort_pairs(nullptr, *bytes, (const uint64_t *)nullptr,
                                   (uint64_t *)nullptr, (const int32_t *)nullptr,
                                   (int32_t *)nullptr, (unsigned)n, 0u,
                                   (unsigned)key_bits, (hipStream_t)0);
}

hipError_t launch_sort_columns(const BuildBuffers &b, const float *xyz,
                               const int64_t *grid, int64_t n,
                               int64_t cells_hint, int depth_bits, int key_bits,
                               hipStream_t s) {
  hipError_t e = launch_count_scan(b, n, cells_hint, s);
  if (e != hipSuccess || n == 0) return e;
  hipLaunchKernelGGL(k_column_keys, dim3(blocks_of(n)), dim3(kBlock), 0, s,
                     b.pt_cell, b.pt_lin, n, b.clouds, depth_bits, b.ckeys[0],
                     b.cperm[0], b.scalars);
  size_t bytes = b.sort_tmp_bytes;
  e = rocprim::radix_sort_pairs(b.sort_tmp, bytes, b.ckeys[0], b.ckeys[1],
                                b.cperm[0], b.cperm[1], (unsigned)n, 0u,
                                (unsigned)key_bits, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scatter_perm, dim3(blocks_of(n)), dim3(kBlock), 0, s, xyz,
                     grid, n, b.cperm[1], b.pt_cell, b.clouds, b.srec, b.scell,
                     b.scalars);
  return hipGetLastError();
}

}  // namespace fk
