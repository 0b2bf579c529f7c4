// Voxel-first subsampling on the device (gfx950).
//
// Semantics of the reference's _voxel_first_cuda
// (Networks/hierarchy.py:23-36): voxel v = floor(x / size) per axis with the
// correctly rounded fp32 quotient (torch's device division by a Python
// scalar); the voxel key is xyz2key of v, which keeps the low 16 bits of
// each axis
// (FlashKNN/functions/z_order.py:56-91); the representative of a voxel is
// the point that comes first after a stable argsort of the keys, i.e. its
// smallest index; the kept indices are returned ascending.  Here each cloud
// of an offset batch is voxelised on its own (the reference passes one
// cloud).
//
// No sort: an open-addressing hash table of 2^s >= 2n slots maps the key
// (cloud, v & 0xffff per axis) to a slot with atomicCAS, atomicMin keeps the
// smallest point index of the slot, and a flag scan compacts the points that
// are their slot's minimum.  The result does not depend on the order in
// which the atomics land.
#include "fk_internal.h"

#include <climits>
#include <mutex>

namespace fk {

hipError_t launch_scan_host(const int32_t *in, int32_t *out, int64_t n,
                            int32_t *part, hipStream_t s);
int64_t scan_parts_needed(int64_t n);

namespace {

constexpr int kBlock = 256;
constexpr uint64_t kEmpty = ~0ull;  // real keys have bit 63 clear

inline unsigned nblocks(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  return (unsigned)(b < 1 ? 1 : b);
}

inline int64_t align_up(int64_t v) { return (v + 255) & ~int64_t(255); }

int64_t table_slots(int64_t n) {
  int64_t s = 64;
  while (s < 2 * n) s <<= 1;
  return s;
}

struct VoxelWs {
  int32_t *status;     // [1]
  uint64_t *keys;      // [S]
  int32_t *mins;       // [S]
  int32_t *pt_slot;    // [n]
  int32_t *flags;      // [n]
  int32_t *pos;        // [n + 1]
  int32_t *part;       // scan partials
  int64_t S;
};

VoxelWs carve(void *ws, int64_t n) {
  VoxelWs w;
  w.S = table_slots(n);
  char *p = (char *)ws;
  w.status = (int32_t *)p; p += 256;
  w.keys = (uint64_t *)p; p += align_up(8 * w.S);
  w.mins = (int32_t *)p; p += align_up(4 * w.S);
  w.pt_slot = (int32_t *)p; p += align_up(4 * n);
  w.flags = (int32_t *)p; p += align_up(4 * n);
  w.pos = (int32_t *)p; p += align_up(4 * (n + 1));
  w.part = (int32_t *)p;
  return w;
}

__global__ void k_vox_init(uint64_t *keys, int32_t *mins, int64_t S,
                           int32_t *status) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < S;
       i += (int64_t)gridDim.x * blockDim.x) {
    keys[i] = kEmpty;
    mins[i] = INT_MAX;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *status = 0;
}

__device__ __forceinline__ uint64_t mix(uint64_t k) {
  // splitmix64 finaliser: spreads neighbouring voxels over the table
  k ^= k >> 30;
  k *= 0xbf58476d1ce4e5b9ull;
  k ^= k >> 27;
  k *= 0x94d049bb133111ebull;
  k ^= k >> 31;
  return k;
}

// Low 16 bits of floor(x / size) (two's complement, as the reference's
// integer voxel coordinates are masked); false outside |v| < 2^31 or for a
// non-finite coordinate.
__device__ __forceinline__ bool voxel16(float x, float size, uint32_t &v16) {
  const float v = floorf(__fdiv_rn(x, size));
  if (!(fabsf(v) < 2147483648.0f)) return false;
  v16 = (uint32_t)(int32_t)v & 0xffffu;
  return true;
}

__global__ void k_vox_insert(const float *__restrict__ xyz,
                             const int32_t *__restrict__ offsets, int32_t B,
                             int32_t n, float size, uint64_t *keys,
                             int32_t *mins, int64_t S, int32_t *pt_slot,
                             int32_t *status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cloud = 0;
  if (offsets) {  // first c with offsets[c] > i (cumulative segment ends)
    int lo = 0, hi = B - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (offsets[mid] > i) hi = mid; else lo = mid + 1;
    }
    cloud = lo;
  }
  uint32_t vx, vy, vz;
  if (!voxel16(xyz[3 * (int64_t)i], size, vx) ||
      !voxel16(xyz[3 * (int64_t)i + 1], size, vy) ||
      !voxel16(xyz[3 * (int64_t)i + 2], size, vz) ||
      (offsets && offsets[cloud] <= i)) {
    atomicCAS(status, 0, (int)FK_ERR_DOMAIN);
    pt_slot[i] = -1;
    return;
  }
  const uint64_t key = ((uint64_t)cloud << 48) | ((uint64_t)vx << 32) |
                       ((uint64_t)vy << 16) | (uint64_t)vz;
  const uint64_t mask = (uint64_t)S - 1;
  uint64_t slot = mix(key) & mask;
  // at most n distinct keys in S >= 2n slots: the probe ends within S steps
  for (int64_t p = 0; p < S; ++p) {
    const uint64_t prev =
        atomicCAS((unsigned long long *)(keys + slot), (unsigned long long)kEmpty,
                  (unsigned long long)key);
    if (prev == kEmpty || prev == key) {
      atomicMin(mins + slot, i);
      pt_slot[i] = (int32_t)slot;
      return;
    }
    slot = (slot + 1) & mask;
  }
  atomicCAS(status, 0, (int)FK_ERR_CAPACITY);  // unreachable for S >= 2n
  pt_slot[i] = -1;
}

__global__ void k_vox_flag(const int32_t *pt_slot, const int32_t *mins,
                           int32_t n, int32_t *flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = pt_slot[i];
  flags[i] = (s >= 0 && mins[s] == i) ? 1 : 0;
}

__global__ void k_vox_compact(const int32_t *flags, const int32_t *pos,
                              int32_t n, const int32_t *status,
                              int64_t *out_idx, int64_t *out_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int st = *status;
  if (i < n && st == 0 && flags[i]) out_idx[pos[i]] = i;
  if (i == 0) *out_count = st != 0 ? -(int64_t)st : (int64_t)pos[n];
}

}  // namespace

// ------------------------------------------------ allocation-free variant
//
// fk_voxel_first_noalloc keeps its workspace in one file-scope device array
// (each GPU holds its own copy, at its own address, until the process ends)
// instead of a caller-provided buffer: for callers that cannot allocate
// (graph capture, serving loops).  Sized for kNoAllocMaxPoints points; the
// host side takes the array's device address per device with
// hipGetSymbolAddress and never passes the array's name to a kernel.
constexpr int64_t kNoAllocMaxPoints = FK_VOXEL_NOALLOC_MAX_POINTS;

constexpr int64_t const_slots(int64_t n) {
  int64_t s = 64;
  while (s < 2 * n) s <<= 1;
  return s;
}
constexpr int64_t const_align(int64_t v) { return (v + 255) & ~int64_t(255); }
// the voxel workspace of kNoAllocMaxPoints points (the layout of carve(),
// scan partials bounded by n / 4096 + 3)
constexpr int64_t kScratchBytes =
    256 + const_align(8 * const_slots(kNoAllocMaxPoints)) +
    const_align(4 * const_slots(kNoAllocMaxPoints)) +
    2 * const_align(4 * kNoAllocMaxPoints) + const_align(4 * (kNoAllocMaxPoints + 1)) +
    const_align(4 * (kNoAllocMaxPoints / 4096 + 3));

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

int64_t voxel_workspace_bytes(int64_t n) {
  const int64_t S = table_slots(n);
  return 256 + align_up(8 * S) + align_up(4 * S) + 2 * align_up(4 * n) +
         align_up(4 * (n + 1)) + align_up(4 * scan_parts_needed(n));
}

hipError_t launch_voxel_first(const float *xyz, const int32_t *offsets,
                              int32_t B, int64_t n, float size, void *ws,
                              int64_t *out_idx, int64_t *out_count,
                              hipStream_t s) {
  VoxelWs w = carve(ws, n);
  const unsigned ib = (unsigned)((w.S + kBlock - 1) / kBlock < 4096
                                     ? (w.S + kBlock - 1) / kBlock
                                     : 4096);
  hipLaunchKernelGGL(k_vox_init, dim3(ib), dim3(kBlock), 0, s, w.keys, w.mins,
                     w.S, w.status);
  if (n > 0) {
    hipLaunchKernelGGL(k_vox_insert, dim3(nblocks(n)), dim3(kBlock), 0, s, xyz,
                       offsets, B, (int32_t)n, size, w.keys, w.mins, w.S,
                       w.pt_slot, w.status);
    hipLaunchKernelGGL(k_vox_flag, dim3(nblocks(n)), dim3(kBlock), 0, s,
                       w.pt_slot, w.mins, (int32_t)n, w.flags);
  }
  // n = 0 writes pos[0] = 0
  hipError_t e = launch_scan_host(w.flags, w.pos, n, w.part, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_vox_compact, dim3(nblocks(n)), dim3(kBlock), 0, s,
                     w.flags, w.pos, (int32_t)n, w.status, out_idx, out_count);
  return hipGetLastError();
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices] = {};  // device address of `scratch` per device
}  // namespace

hipError_t launch_voxel_first_noalloc(const float *xyz, const int32_t *offsets,
                                      int32_t B, int64_t n, float size,
                                      int64_t *out_idx, int64_t *out_count,
                                      hipStream_t s) {
  if (n > kNoAllocMaxPoints || voxel_workspace_bytes(n) > kScratchBytes)
    return hipErrorInvalidValue;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  void *ws = nullptr;
  {
    // first call on this device (it is current): look the array up once
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess) return e;
      g_scratch_addr[dev] = p;
    }
    ws = g_scratch_addr[dev];
  }
  // every workspace word the kernels read is written first in the same call
  // (hash table and status by k_vox_init, flags, scan partials and
  // positions by their producers), so the previous call's contents are
  // harmless
  return launch_voxel_first(xyz, offsets, B, n, size, ws, out_idx, out_count, s);
}

}  // namespace fk
