// Grid kNN query kernels for gfx950 (wave64).
//
// Semantics follow FlashKNN_Query_dynamic_load_kernel
// (FlashKNN/csrc/flash_knn_query_dynamic_load.cu:469-652): candidates are
// the support points of the query's cloud in the cells within R of the
// query's cell (R = 1: 27 cells), d = ((dx*dx + dy*dy) + dz*dz) with
// dx = p - q (:601-604, evaluated here without contraction), a candidate
// enters only if d < cut_r2 (:606), slots without a candidate hold
// (pad, +inf) (:562), rows are written ascending.  Ties are broken by the
// smaller support id (the reference keeps the first one scanned).
//
// Mapping (designed for 64-lane wavefronts, not a translation of the
// reference's lane groups): a group of G lanes (1, 2 or 4 by K class) owns
// one query and a wave owns 64 / G consecutive queries of the cell-sorted
// order, so its lanes walk the same or adjacent column segments and their
// candidate loads coalesce or broadcast; lane g of a group takes every G-th
// record of each column.  Each lane keeps its best KC keys sorted in VGPRs
// (keys below: one compare orders distance and tie together).  Accepted
// candidates are appended to a per-lane LDS buffer (one ds_write, no
// divergent insertion); when any lane's buffer is nearly full the wave runs
// one bitonic sort of the buffer and a bitonic merge into the register
// list, and the group adopts its smallest acceptance threshold.  At the end
// the group's lists are merged by a butterfly over LDS.  More lanes per
// query shorten each wave's chain of dependent loads, which bounds the pass
// (a 100k-point cloud is only 1,536 waves of 64 queries).
//
// FK_MODE_EXACT adds a certification per query: if the K-th distance is
// below the distance from the query to the nearest face of its scanned
// block (faces beyond which the cloud has no cells do not count), no point
// outside the block can enter, and the result equals brute force.  The rest
// go to a pending list and a second kernel grows their block shell by shell
// (radius from the current K-th distance; while there is none, from a probe
// of full-height block counts, else grown geometrically) until they
// certify.
//
// The file is compiled twice: with FK_COLUMNS=0 for grids with a dense cell
// table, and with FK_COLUMNS=1 for the column layout (column_range searches
// z inside a column).  Only the lookup differs; keeping it out of the dense
// kernels keeps their register allocation.  The public launchers of the
// column build carry a _columns suffix, and the dense ones forward to them
// when the grid uses the column layout.
#include "fk_internal.h"

#include <cfloat>
#include <climits>
#include <type_traits>

#ifndef FK_COLUMNS
#define FK_COLUMNS 0
#endif
#if FK_COLUMNS
#define FK_LAYOUT(name) name##_columns
#else
#define FK_LAYOUT(name) name
#endif

// Translation units: the file is compiled once per part and layout, so the
// parts build in parallel: FK_QPART = 8, 16, 32 or 64 holds that K class's
// first-pass and batched kernels and their launchers, 2 the cooperative
// expansion, 3 the one-wave-per-query kernel of k > 64, and 1 the rest
// (counts, the plan, work items, the dispatching launchers); FK_QPART = 0
// compiles all of it as one unit.
#ifndef FK_QPART
#define FK_QPART 0
#endif
#define FK_COMMON (FK_QPART == 0 || FK_QPART == 1)
#define FK_EXPAND_PART (FK_QPART == 0 || FK_QPART == 2)
#define FK_WIDE_PART (FK_QPART == 0 || FK_QPART == 3)
#define FK_CLASS(kc) (FK_QPART == 0 || FK_QPART == (kc))
// Each part (and the column layout) uses only some of the shared helpers.
#pragma clang diagnostic ignored "-Wunused-function"
#pragma clang diagnostic ignored "-Wunused-const-variable"
#pragma clang diagnostic ignored "-Wunneeded-internal-declaration"

namespace fk {

namespace {

// Tuning knobs (compile time; the defaults are the measured best):
//   FK_UNROLL        candidates loaded per lane per step
//   FK_HALF_MASK     bit 0/1/2: K class 16/32/64 buffers KC/2 keys (merged
//                    with a half-cleaner) instead of KC
//   FK_WPE_<KC>      minimum waves per SIMD requested for the first pass
//                    (3 for K class 16 holds it to 168 VGPRs, three waves
//                    per SIMD: the planned kernel fits, the uniform one
//                    spills a few bytes; faster than 171 VGPRs at two)
//   FK_SLOTS         records per lane per round of the K class 8/16 stream
//   FK_FILL_WAVES    a launch with fewer waves widens its lane groups
#ifndef FK_UNROLL
#define FK_UNROLL 4
#endif
#ifndef FK_HALF_MASK
#define FK_HALF_MASK 0
#endif
#ifndef FK_WPE_8
#define FK_WPE_8 1
#endif
#ifndef FK_WPE_16
#define FK_WPE_16 3
#endif
#ifndef FK_WPE_32
#define FK_WPE_32 2
#endif
#ifndef FK_WPE_64
#define FK_WPE_64 1
#endif
//   FK_G_<KC>        lanes per query in the first pass (1, 2 or 4); K class
//                    32 takes one (250k-point room, a launch of many more
//                    tiles than wave slots: exact 201 -> 187 us, 27-cell
//                    188 -> 171 us; K class 16 stays at two, 96 vs 121 us)
//   FK_INLINE_PLAIN  the unplanned first pass (k_query) expands a tile's
//                    lone uncertified query itself, as the planned one does
#ifndef FK_INLINE_PLAIN
#define FK_INLINE_PLAIN 1
#endif
#ifndef FK_G_8
#define FK_G_8 2
#endif
#ifndef FK_G_16
#define FK_G_16 2
#endif
#ifndef FK_G_32
#define FK_G_32 1
#endif
#ifndef FK_G_64
#define FK_G_64 1
#endif

constexpr int kUnroll = FK_UNROLL;
// Keys: ((d bits << 32) | id) + 2^52.  For d >= 0 the u64 order is the
// (d, id) order, and with the bias every key is a positive normal double
// whose double order is the same, so compare-exchanges use v_min_f64 /
// v_max_f64 (which return an input exactly) instead of a 64-bit compare and
// four selects.  kMaxKey is +inf as a double, above every key.
constexpr uint64_t kKeyBias = 1ull << 52;
constexpr uint64_t kMaxKey = 0x7ff0000000000000ull;

template <int KC>
constexpr int buffer_slots() {
  // K class 8 buffers 16 keys so that it can take the 8-stream rounds
  return KC == 8 ? 16
         : ((KC == 16 && (FK_HALF_MASK & 1)) || (KC == 32 && (FK_HALF_MASK & 2)) ||
            (KC == 64 && (FK_HALF_MASK & 4)))
             ? KC / 2
             : KC;
}

template <int KC>
constexpr int lanes_per_query() {
  return KC == 8 ? FK_G_8 : KC == 16 ? FK_G_16 : KC == 32 ? FK_G_32 : FK_G_64;
}

template <int KC>
constexpr int min_waves_per_eu() {
  return KC == 8 ? FK_WPE_8 : KC == 16 ? FK_WPE_16 : KC == 32 ? FK_WPE_32 : FK_WPE_64;
}
constexpr uint32_t kInfBits = 0x7f800000u;

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

// LDS writes of the wave visible to its own later LDS reads.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// The value lane (lane ^ J) holds, J in {1, 2, 4, 8, 16, 32}, without a
// round trip through LDS (__shfl_xor compiles to ds_bpermute): DPP inside a
// row of 16 lanes (quad_perm for 1 and 2, row_half_mirror then a reversed
// quad for 4, row_ror:8 for 8) and the gfx950 row swaps v_permlane16_swap /
// v_permlane32_swap for 16 and 32.  Every lane of the wave must be active.
template <int J>
__device__ __forceinline__ uint32_t lane_xor32(uint32_t v) {
  if constexpr (J == 1) {
    return __builtin_amdgcn_update_dpp(0u, v, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
  } else if constexpr (J == 2) {
    return __builtin_amdgcn_update_dpp(0u, v, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
  } else if constexpr (J == 4) {
    const uint32_t h = __builtin_amdgcn_update_dpp(0u, v, 0x141, 0xF, 0xF, false);
    return __builtin_amdgcn_update_dpp(0u, h, 0x1B, 0xF, 0xF, false);  // quad_perm 3,2,1,0
  } else if constexpr (J == 8) {
    return __builtin_amdgcn_update_dpp(0u, v, 0x128, 0xF, 0xF, false);  // row_ror:8
  } else if constexpr (J == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (threadIdx.x & 16) ? r[0] : r[1];
  } else {
    static_assert(J == 32, "lane_xor: J in {1, 2, 4, 8, 16, 32}");
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (threadIdx.x & 32) ? r[0] : r[1];
  }
}

template <int J>
__device__ __forceinline__ uint64_t lane_xor(uint64_t v) {
  return ((uint64_t)lane_xor32<J>((uint32_t)(v >> 32)) << 32) | lane_xor32<J>((uint32_t)v);
}

// lane_xor with the distance known after unrolling.
__device__ __forceinline__ uint64_t lane_xor_n(uint64_t v, int j) {
  switch (j) {
    case 1: return lane_xor<1>(v);
    case 2: return lane_xor<2>(v);
    case 4: return lane_xor<4>(v);
    case 8: return lane_xor<8>(v);
    case 16: return lane_xor<16>(v);
    default: return lane_xor<32>(v);
  }
}

// The value lane 63 - lane holds: row_mirror inside each row of 16, then
// the rows in reverse order (row r <- row 3 - r: xor 16, then xor 32).
__device__ __forceinline__ uint64_t lane_reverse(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_update_dpp(0u, (uint32_t)v, 0x140, 0xF, 0xF, false);
  const uint32_t hi = __builtin_amdgcn_update_dpp(0u, (uint32_t)(v >> 32), 0x140, 0xF, 0xF, false);
  return lane_xor<32>(lane_xor<16>(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ float sqdist(float px, float py, float pz, float qx,
                                        float qy, float qz) {
  const float dx = __fsub_rn(px, qx);
  const float dy = __fsub_rn(py, qy);
  const float dz = __fsub_rn(pz, qz);
  return __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)),
                   __fmul_rn(dz, dz));
}

__device__ __forceinline__ uint64_t key_of_bits(uint32_t dbits, uint32_t id) {
  return (((uint64_t)dbits << 32) | id) + kKeyBias;
}

__device__ __forceinline__ uint64_t make_key(float d, uint32_t id) {
  return key_of_bits(__float_as_uint(d), id);
}

__device__ __forceinline__ uint32_t key_dbits(uint64_t key) {
  return (uint32_t)((key - kKeyBias) >> 32);
}

__device__ __forceinline__ uint32_t key_id(uint64_t key) { return (uint32_t)key; }

// Keys are never NaN, so the min/max are issued directly (fmin/fmax would
// add a NaN-quieting v_max_f64 per input in IEEE mode).
__device__ __forceinline__ uint64_t kmin(uint64_t a, uint64_t b) {
  double r;
  asm("v_min_f64 %0, %1, %2"
      : "=v"(r)
      : "v"(__longlong_as_double((long long)a)), "v"(__longlong_as_double((long long)b)));
  return (uint64_t)__double_as_longlong(r);
}

__device__ __forceinline__ uint64_t kmax(uint64_t a, uint64_t b) {
  double r;
  asm("v_max_f64 %0, %1, %2"
      : "=v"(r)
      : "v"(__longlong_as_double((long long)a)), "v"(__longlong_as_double((long long)b)));
  return (uint64_t)__double_as_longlong(r);
}

__device__ __forceinline__ void cswap(uint64_t &a, uint64_t &b, bool asc) {
  const uint64_t lo = kmin(a, b), hi = kmax(a, b);
  a = asc ? lo : hi;
  b = asc ? hi : lo;
}

// Ascending bitonic sort of N (power of two) register keys.
template <int N>
__device__ __forceinline__ void bitonic_sort(uint64_t (&a)[N]) {
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int l = i ^ j;
        if (l > i) cswap(a[i], a[l], (i & k) == 0);
      }
    }
  }
}

// Bitonic input -> ascending.
template <int N>
__device__ __forceinline__ void bitonic_merge(uint64_t (&a)[N]) {
#pragma unroll
  for (int j = N >> 1; j > 0; j >>= 1) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int l = i ^ j;
      if (l > i) cswap(a[i], a[l], true);
    }
  }
}

// Per-lane running top-k (sorted ascending) + LDS append buffer.
//
// The k real entries live in the upper k slots of a KC-slot register list;
// slots [0, KC-k) hold key 0, which sorts below every real key and so stays
// there through every merge.  The k-th best is then always top[KC-1]: no
// register is ever indexed by the runtime k (a runtime index would force the
// whole list into scratch memory).
template <int KC>
struct TopK {
  uint64_t top[KC];
  uint64_t *lb;     // this lane's buffer: slot s at lb[s * 64]
  int cnt;
  int k;
  uint32_t cutbits;  // bits of the cut distance
  uint32_t pad;      // index of an empty slot
  uint64_t thr;      // accept key < thr
  // (32-bit fields and derived values: the K class 16 pass runs at the edge
  // of three waves per SIMD, where every VGPR counts)

  __device__ __forceinline__ int base() const { return KC - k; }  // first real slot
  __device__ __forceinline__ uint64_t cutkey() const { return key_of_bits(cutbits, 0u); }
  __device__ __forceinline__ uint64_t padkey() const { return key_of_bits(kInfBits, pad); }

  __device__ __forceinline__ void init(uint64_t *lane_buf, int k_, float cut_r2,
                                       uint32_t pad_) {
    lb = lane_buf;
    cnt = 0;
    k = k_;
    cutbits = __float_as_uint(cut_r2);
    pad = pad_;
#pragma unroll
    for (int i = 0; i < KC; ++i) top[i] = i < base() ? 0ull : padkey();
    thr = cutkey();  // (inf, pad) >= cutkey for every cut
  }

  __device__ __forceinline__ void refresh_thr() {
    const uint64_t ck = cutkey();
    thr = top[KC - 1] < ck ? top[KC - 1] : ck;
  }

  // Merge the buffered keys into the sorted list (uniform work per lane).
  __device__ __forceinline__ void flush() {
    constexpr int B = buffer_slots<KC>();
    uint64_t b[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const uint64_t v = lb[i * kWave];
      b[i] = i < cnt ? v : kMaxKey;
    }
    bitonic_sort<B>(b);
    if constexpr (B >= KC) {
      // min(top[i], b[KC-1-i]) is bitonic and holds the KC smallest (only
      // the KC smallest buffered keys can enter)
#pragma unroll
      for (int i = 0; i < KC; ++i) top[i] = kmin(top[i], b[KC - 1 - i]);
      bitonic_merge<KC>(top);
    } else {
      // B = KC/2: the half-cleaner min(top[B+i], b[B-1-i]) gives the B
      // smallest of the upper half and the buffer (bitonic); sort it, then
      // merge the two sorted halves (second half read reversed = bitonic).
      uint64_t h[B];
#pragma unroll
      for (int i = 0; i < B; ++i) h[i] = kmin(top[B + i], b[B - 1 - i]);
      bitonic_merge<B>(h);
#pragma unroll
      for (int i = 0; i < B; ++i) top[KC - 1 - i] = h[i];
      bitonic_merge<KC>(top);
    }
    cnt = 0;
    refresh_thr();
  }

  // Append key if ok and below the threshold.  The store is unconditional
  // (slot cnt is free: a rejected key is overwritten by the next append), so
  // the append is a store and an add instead of a masked branch; callers
  // guarantee cnt < buffer_slots<KC>() at every call.
  __device__ __forceinline__ void push(uint64_t key, bool ok) {
    lb[cnt * kWave] = key;
    cnt += (ok && key < thr) ? 1 : 0;
  }

  __device__ __forceinline__ uint64_t kth() const { return top[KC - 1]; }

  // G lanes scan disjoint parts of one query's candidates.  The k-th key of
  // their union is <= each lane's k-th, so every lane may reject against the
  // smallest threshold of its group.
  template <int G>
  __device__ __forceinline__ void share_threshold() {
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
      const uint64_t u = lane_xor_n(thr, o);
      thr = u < thr ? u : thr;
    }
  }

  // Butterfly merge of the group's partial lists: afterwards every lane of
  // the group holds the k best keys of the union.  Each round a lane copies
  // its k real keys to slots 0..k-1 of its (empty) LDS buffer and reads its
  // partner's; min(top[i], partner[KC-1-i]) is bitonic over the real slots
  // (the zero slots below `base` meet +inf), and one bitonic merge sorts it.
  template <int G>
  __device__ __forceinline__ void group_merge(int lane) {
    static_assert(G == 1 || buffer_slots<KC>() >= KC, "merge needs k buffer slots");
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
#pragma unroll
      for (int j = 0; j < KC; ++j)
        if (j >= base()) lb[(j - base()) * kWave] = top[j];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // partner key j (j < k) is read where it is used: no second list
      const uint64_t *pb = lb + ((lane ^ o) - lane);
#pragma unroll
      for (int i = 0; i < KC; ++i) {
        const int j = KC - 1 - i;
        top[i] = kmin(top[i], j < k ? pb[j * kWave] : kMaxKey);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      bitonic_merge<KC>(top);
    }
    refresh_thr();
  }

  // Mid-scan merge: the group's lead lane keeps the merged list, the others
  // start again from empty slots (so the final merge counts nothing twice),
  // and every lane takes the merged k-th key as its threshold.
  template <int G>
  __device__ __forceinline__ void group_merge_keep(int lane, int sub) {
    group_merge<G>(lane);
    if (sub != 0) {
#pragma unroll
      for (int i = 0; i < KC; ++i) top[i] = i < base() ? 0ull : padkey();
      thr = cutkey();
    }
    share_threshold<G>();
  }
};

// Scan support records [b, e) for this lane's query.  Wave-uniform trip
// count (the longest lane); a merge runs whenever any lane could overflow.
//
// The rounds that only append (the inner loop) never touch the register
// list: it changes only in the flush between them, so the compiler keeps it
// in place instead of copying it around every round.
//
// With G lanes per query, lane `sub` of the group takes records
// b + sub, b + sub + G, ... (the group's loads are adjacent).
template <int KC, int G, int U = kUnroll>
__device__ __forceinline__ void scan_range(TopK<KC> &t, int b, int e, int sub,
                                           float qx, float qy, float qz,
                                           const float4 *__restrict__ srec) {
  constexpr int kRoom = buffer_slots<KC>() - U;
  static_assert(kRoom >= 0, "buffer too small for a round");
  const int span = e - b - sub;
  const int len = span > 0 ? (span + G - 1) / G : 0;  // this lane's records
  const int first = b + sub;
  int s = 0;
  for (;;) {
    while (wave_any(s < len) && !wave_any(t.cnt > kRoom)) {
      float4 r[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = s + u < len;
        r[u] = srec[ok ? (uint32_t)(first + (s + u) * G) : 0u];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float d = sqdist(r[u].x, r[u].y, r[u].z, qx, qy, qz);
        t.push(make_key(d, __float_as_uint(r[u].w)), s + u < len);
      }
      s += U;
    }
    if (!wave_any(s < len)) break;
    t.flush();
    t.template share_threshold<G>();
  }
}

// Column segment (cx, cy, z0..z1) of the query's cloud box -> [b, e).
// Dense boxes read the two ends from the cell table; a column layout reads
// the column's run and narrows it to z0..z1 by binary search over the
// sorted local z of its records.
__device__ __forceinline__ void column_range(const CloudInfo &ci,
                                             const int32_t *__restrict__ cs,
                                             const int4 *__restrict__ sc,
                                             int cx, int cy, int z0, int z1,
                                             int &b, int &e) {
  b = e = 0;
  if (cx < 0 || cx >= ci.nx || cy < 0 || cy >= ci.ny) return;
  z0 = z0 < 0 ? 0 : z0;
  z1 = z1 >= ci.nz ? ci.nz - 1 : z1;
  if (z0 > z1) return;
  if constexpr (!FK_COLUMNS) {
    (void)sc;
    const int col = ci.base + (cx * ci.ny + cy) * ci.nz;
    b = cs[col + z0];
    e = cs[col + z1 + 1];
  } else {
    const int col = ci.base + cx * ci.ny + cy;
    int lo = cs[col];
    const int hi = cs[col + 1];
    int h = hi;
    while (lo < h) {  // first record with z >= z0
      const int m = lo + ((h - lo) >> 1);
      if (sc[m].z < z0) lo = m + 1; else h = m;
    }
    b = lo;
    h = hi;
    while (lo < h) {  // first record with z > z1
      const int m = lo + ((h - lo) >> 1);
      if (sc[m].z <= z1) lo = m + 1; else h = m;
    }
    e = lo;
  }
}

// Scan the full block of half-width R around the query's cell.  The
// query's own column goes first: its points are the nearest on average, so
// the k-th distance (the acceptance threshold) is tight early and fewer
// later candidates get buffered and merged.  Then the others in raster
// order (dx, dy) = (0, 0) is skipped.
// Exact pruning of neighbour cells against the current threshold (grids
// built from coordinates, where a cell's extent is known).  In cell units
// (x * inv_cell), a point of the cell at offset d on an axis lies at least
// gap(d) from the query on that axis: f or 1 - f plus |d| - 1 whole cells,
// f the query's position inside its own cell.  The margin m covers the
// fp32 rounding of x * inv (which decides a point's cell) and of the
// query's own position; the 1e-5 factors cover the fp32 distance and
// threshold arithmetic.  A cell is skipped only if every point in it is
// strictly farther than the threshold, so the result does not change.
struct Cull {
  float f[3], m[3];
  float eff_s;  // threshold distance in cell units^2, inflated; +inf: no pruning
  __device__ __forceinline__ void setup(const CloudInfo &ci, int4 qc, float4 qr,
                                        float eff) {
    const float inv = ci.inv_cell;
    const float u[3] = {__fmul_rn(qr.x, inv), __fmul_rn(qr.y, inv),
                        __fmul_rn(qr.z, inv)};
    const int c[3] = {qc.x + ci.minx, qc.y + ci.miny, qc.z + ci.minz};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      f[i] = u[i] - (float)c[i];
      m[i] = 4e-7f * (fabsf(u[i]) + 4.f);
    }
    eff_s = (inv > 0.f && eff < INFINITY)
                ? __fmul_rn(__fmul_rn(__fmul_rn(eff, inv), inv), 1.00001f)
                : INFINITY;
  }
  __device__ __forceinline__ float gap2(int ax, int d) const {
    float g = d < 0 ? f[ax] + (float)(-d - 1) : (d > 0 ? (1.f - f[ax]) + (float)(d - 1) : 0.f);
    g = g > m[ax] ? g - m[ax] : 0.f;
    return g * g;
  }
  __device__ __forceinline__ bool keep(float lb2) const {
    return __fmul_rn(lb2, 0.99999f) <= eff_s;
  }
};

// Column (qc.x + dx, qc.y + dy), z within qc.z -/+ 1, trimmed to the cells
// the threshold can still reach: [b, e) (empty when the column is pruned).
__device__ __forceinline__ void culled_column(const CloudInfo &ci,
                                              const int32_t *__restrict__ cs,
                                              const int4 *__restrict__ sc,
                                              const Cull &cu, int4 qc, int dx,
                                              int dy, int &b, int &e) {
  b = e = 0;
  const float xy = cu.gap2(0, dx) + cu.gap2(1, dy);
  if (!cu.keep(xy)) return;
  const int z0 = cu.keep(xy + cu.gap2(2, -1)) ? qc.z - 1 : qc.z;
  const int z1 = cu.keep(xy + cu.gap2(2, 1)) ? qc.z + 1 : qc.z;
  column_range(ci, cs, sc, qc.x + dx, qc.y + dy, z0, z1, b, e);
}

// The candidate ranges of one query (its own column, then the 8 around it)
// as one stream: lane `sub` of the query's group takes every G-th record of
// each range, the ranges are consumed one after another, kSlots consecutive
// records per round, and a round runs on into the next range when the
// current one ends.  A lane's rounds then follow its total record count;
// eight parallel column streams followed the longest of its columns (about
// twice the mean on indoor scans, several times on LiDAR rings), and the
// rounds of a wave are the most any of its lanes needs.
#ifndef FK_SLOTS
#define FK_SLOTS 8
#endif
constexpr int kSlots = FK_SLOTS;
constexpr int kStreamCols = 9;  // LDS table entries per lane

// Candidate batch: the records a lane loads per round before the wave
// checks whether any lane's key buffer could overflow.  A longer batch keeps
// more loads in flight and checks less often, at 4 VGPRs per record and
// less room left in the buffer (more merges); the best length depends on
// the cloud (candidates per query, acceptance rate), so the first pass is
// compiled for three: the K classes 8/16 take it in the stream scan
// (default FK_SLOTS), 32/64 in the column scan (default FK_UNROLL), and
// FK_QUERY_BATCH_SMALL / _LARGE select the alternatives (results never
// change).
enum BatchSel : int { kBatchDefault = 0, kBatchSmall = 1, kBatchLarge = 2 };
template <int KC>
constexpr int batch_records(int sel) {
  return KC <= 16 ? (sel == kBatchSmall ? kSlots / 2
                     : sel == kBatchLarge ? buffer_slots<KC>() * 3 / 4 : kSlots)
                  : (sel == kBatchSmall ? kUnroll / 2
                     : sel == kBatchLarge ? 2 * kUnroll : kUnroll);
}
template <int KC>
constexpr int default_batch() { return batch_records<KC>(kBatchDefault); }
// The column scan's batch in a kernel of candidate batch CB: CB for the K
// classes that scan columns, the default for the stream classes (their
// column scans serve only blocks wider than 27 cells).
template <int KC, int CB>
constexpr int range_batch() { return KC <= 16 ? kUnroll : CB; }

// A record range [b, e) in one 64-bit LDS word of a key buffer, moved by
// byte copies (so the words may hold keys at other times without any
// type-based aliasing assumption).
__device__ __forceinline__ void put_range(uint64_t *slot, int b, int e) {
  const int2 r = make_int2(b, e);
  __builtin_memcpy(slot, &r, sizeof(r));
}
__device__ __forceinline__ int2 get_range(const uint64_t *slot) {
  int2 r;
  __builtin_memcpy(&r, slot, sizeof(r));
  return r;
}

// Table entries a kernel of K class KC needs (the stream scan serves K
// classes 8 and 16; the others keep 1 so that the array is never empty).
template <int KC>
constexpr int stream_entries() {
  return KC <= 16 ? kStreamCols : 1;
}

// The ranges live in a lane-private LDS table (entry c at tab[c * kWave]):
// in registers they would cost 18 VGPRs and a wave per SIMD.
template <int NC, int G, int S = kSlots>
struct ColumnStream {
  uint64_t *tab;  // this lane's table: (first record of this lane, range end)
  uint32_t rest;  // non-empty ranges not yet started (bit c)
  int pc, lc;     // current range: next record, records left
  int pn, ln;     // the following non-empty range (ln = 0: none)

  __device__ __forceinline__ void init(uint64_t *lane_tab, const int (&b)[NC],
                                       const int (&en)[NC], int sub) {
    static_assert(NC <= kStreamCols, "stream table too small");
    tab = lane_tab;
    rest = 0u;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int s = b[c] + sub;
      put_range(tab + c * kWave, s, en[c]);
      if (s < en[c]) rest |= 1u << c;
    }
    pop(pc, lc);
    pop(pn, ln);
  }

  // Next non-empty range: its first record and record count (0 when done).
  __device__ __forceinline__ void pop(int &p, int &l) {
    p = 0;
    l = 0;
    if (rest) {
      const int c = __builtin_ctz(rest);
      rest &= rest - 1u;
      const int2 r = get_range(tab + c * kWave);
      p = r.x;
      l = (r.y - r.x + G - 1) / G;
    }
  }
  // Record of slot u of this round: the current range's u-th, or the
  // (u - lc)-th of the following one; index 0 for a slot past both (its
  // load is discarded).  One select per slot: pn + (u - lc) G is
  // (pn - lc G) + u G.
  __device__ __forceinline__ uint32_t addr(int u, bool &ok) const {
    ok = u < lc + ln;
    const int base = u < lc ? pc : pn - lc * G;
    return ok ? (uint32_t)(base + u * G) : 0u;
  }

  // After a round of S records: the current range continues, or the
  // following one (partly taken) becomes current, or both are done.
  __device__ __forceinline__ void advance() {
    if (lc > S) {
      pc += S * G;
      lc -= S;
      return;
    }
    const int used = S - lc;
    if (ln > used) {
      pc = pn + used * G;
      lc = ln - used;
      pop(pn, ln);
    } else {
      pop(pc, lc);
      pop(pn, ln);
    }
  }
};

// GT: the lanes that share one query's threshold (G, unless the group's
// lanes scan different ranges whole).
template <int KC, int G, int NC, int S, int GT = G>
__device__ __forceinline__ void scan_stream(TopK<KC> &t, ColumnStream<NC, G, S> &cs,
                                            float4 qr,
                                            const float4 *__restrict__ srec) {
  constexpr int kRoom = buffer_slots<KC>() - S;
  static_assert(kRoom >= 0, "buffer too small for a round");
  for (;;) {
    while (wave_any(cs.lc > 0) && !wave_any(t.cnt > kRoom)) {
      float4 r[S];
      bool ok[S];
#pragma unroll
      for (int u = 0; u < S; ++u) {
        r[u] = srec[cs.addr(u, ok[u])];
      }
#pragma unroll
      for (int u = 0; u < S; ++u) {
        const float d = sqdist(r[u].x, r[u].y, r[u].z, qr.x, qr.y, qr.z);
        t.push(make_key(d, __float_as_uint(r[u].w)), ok[u]);
      }
      cs.advance();
    }
    if (!wave_any(cs.lc > 0)) break;
    t.flush();
    t.template share_threshold<GT>();
  }
}

// The 8 columns around the query's own, z within qc.z -/+ 1, trimmed by
// the cull (all of them when it does not prune), into slots [o, o + 8).
template <int NC>
__device__ __forceinline__ void ring_columns(const QueryArgs &a, const CloudInfo &ci,
                                             int4 qc, bool active, const Cull &cu,
                                             int (&b)[NC], int (&e)[NC], int o) {
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int j = c < 4 ? c : c + 1;  // 3x3 raster order without the centre
    b[o + c] = e[o + c] = 0;
    if (active) culled_column(ci, a.cell_start, a.scell, cu, qc, j / 3 - 1, j % 3 - 1, b[o + c], e[o + c]);
  }
}

template <int KC, int G, int CB>
__device__ __forceinline__ void scan_block(TopK<KC> &t, const QueryArgs &a,
                                           const CloudInfo &ci, int4 qc,
                                           float4 qr, int R, bool active,
                                           int sub, int lane, uint64_t *lane_tab) {
  // K classes 8 and 16, R = 1: the query's own column, then the 8 neighbour
  // columns, as one stream of records.  With FK_QUERY_PRUNE an exact threshold for
  // the group after the own column (flush + mid-scan merge) prunes the
  // neighbour columns and their end cells; it pays where the 27-cell block
  // holds many more candidates than k (large cells, uneven density).  (The
  // larger classes keep the plain column order: their register lists leave
  // no room for the extra state.)
  if constexpr (KC <= 16 && buffer_slots<KC>() >= 16) {
    if (R == 1) {
      int b0 = 0, e0 = 0;
      if (active)
        column_range(ci, a.cell_start, a.scell, qc.x, qc.y, qc.z - 1, qc.z + 1, b0, e0);
      if (a.prune) {  // launch-uniform
        {
          const int b[1] = {b0}, e[1] = {e0};
          ColumnStream<1, G, CB> cs;
          cs.init(lane_tab, b, e, sub);
          scan_stream<KC, G, 1, CB>(t, cs, qr, a.srec);
        }
        if (wave_any(t.cnt > 0)) t.flush();
        if constexpr (G > 1) t.template group_merge_keep<G>(lane, sub);
        Cull cu;
        cu.setup(ci, qc, qr, __uint_as_float(key_dbits(t.thr)));
        int b[8], e[8];
        ring_columns<8>(a, ci, qc, active, cu, b, e, 0);
        ColumnStream<8, G, CB> cs;
        cs.init(lane_tab, b, e, sub);
        scan_stream<KC, G, 8, CB>(t, cs, qr, a.srec);
      } else {
        Cull cu;
        cu.setup(ci, qc, qr, INFINITY);
        int b[9], e[9];
        b[0] = b0;
        e[0] = e0;
        ring_columns<9>(a, ci, qc, active, cu, b, e, 1);
        ColumnStream<9, G, CB> cs;
        cs.init(lane_tab, b, e, sub);
        scan_stream<KC, G, 9, CB>(t, cs, qr, a.srec);
      }
      return;
    }
  }
  {
    int b = 0, e = 0;
    if (active)
      column_range(ci, a.cell_start, a.scell, qc.x, qc.y, qc.z - R, qc.z + R, b, e);
    scan_range<KC, G, range_batch<KC, CB>()>(t, b, e, sub, qr.x, qr.y, qr.z, a.srec);
  }
  for (int dx = -R; dx <= R; ++dx) {
    for (int dy = -R; dy <= R; ++dy) {
      if (dx == 0 && dy == 0) continue;
      int b = 0, e = 0;
      if (active)
        column_range(ci, a.cell_start, a.scell, qc.x + dx, qc.y + dy, qc.z - R,
                     qc.z + R, b, e);
      scan_range<KC, G, range_batch<KC, CB>()>(t, b, e, sub, qr.x, qr.y, qr.z, a.srec);
    }
  }
}

// Distance (scaled cell units) from the query to the nearest face of its
// block beyond which the cloud's box still has cells; +inf when the block
// covers the box.
__device__ __forceinline__ double cert_gap(const CloudInfo &ci, int4 qc,
                                           float4 qr, int R, double &umax) {
  const float inv = ci.inv_cell;
  const float ux = __fmul_rn(qr.x, inv), uy = __fmul_rn(qr.y, inv),
              uz = __fmul_rn(qr.z, inv);
  umax = fmax(fabs((double)ux), fmax(fabs((double)uy), fabs((double)uz)));
  double gap = HUGE_VAL;
  const double f[3] = {(double)ux - (double)(qc.x + ci.minx),
                       (double)uy - (double)(qc.y + ci.miny),
                       (double)uz - (double)(qc.z + ci.minz)};
  const int c[3] = {qc.x, qc.y, qc.z};
  const int n[3] = {ci.nx, ci.ny, ci.nz};
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    if (c[ax] - R > 0) gap = fmin(gap, f[ax] + R);
    if (c[ax] + R < n[ax] - 1) gap = fmin(gap, (double)(R + 1) - f[ax]);
  }
  return gap;
}

// eff = min(K-th distance, cut): every point with d < eff could still enter.
__device__ __forceinline__ bool certified(double gap, float eff, float inv,
                                          double umax, int R) {
  if (gap == HUGE_VAL) return true;
  if (!(eff < INFINITY)) return false;
  const double g = gap - 2.4e-7 * (umax + R + 2.0);
  if (g <= 0.0) return false;
  const double s = (double)eff * (double)inv * (double)inv * (1.0 + 2e-6);
  return s < g * g;
}

// Smallest block half-width that can certify `eff`, at least Ro + 1.
// Without a k-th distance yet (found < k), the block grows by the factor
// that would hold k points if their count grew like the area of a surface,
// sqrt(k / found), between 1.25 and the doubling 2 + 1/Ro (found = 0:
// doubling): scanned-out sparse surroundings (LiDAR's far range) would
// otherwise double into shells of thousands of empty columns.
// Shell growth while fewer than k candidates are known: by sqrt(k / found)
// (points on surfaces: counts grow with the area), clamped to
// [1.25, FK_GROW_MAX], and FK_GROW_EMPTY * R + 1 with none found.
#ifndef FK_GROW_MAX
#define FK_GROW_MAX 2.f
#endif
#ifndef FK_GROW_EMPTY
#define FK_GROW_EMPTY 2
#endif

__device__ __forceinline__ int next_radius(const CloudInfo &ci, int4 qc,
                                           float4 qr, int Ro, float eff,
                                           int Rcover, int found, int k) {
  int Rn;
  if (eff < INFINITY) {
    const float inv = ci.inv_cell;
    const float u[3] = {__fmul_rn(qr.x, inv), __fmul_rn(qr.y, inv),
                        __fmul_rn(qr.z, inv)};
    const int c[3] = {qc.x + ci.minx, qc.y + ci.miny, qc.z + ci.minz};
    const double rho = sqrt((double)eff * (1.0 + 4e-6)) * (double)inv;
    double need = (double)(Ro + 1);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const double f = (double)u[ax] - (double)c[ax];
      const double m = 2.4e-7 * (fabs((double)u[ax]) + Rcover + 2.0) + 1e-9;
      need = fmax(need, ceil(rho + m - f));
      need = fmax(need, ceil(rho + m - 1.0 + f));
    }
    Rn = need > (double)Rcover ? Rcover : (int)need;
  } else {
    long long r2 = (long long)FK_GROW_EMPTY * Ro + 1;
    if (found > 0 && Ro > 0) {
      const float g = fminf(fmaxf(sqrtf((float)k / (float)found), 1.25f), FK_GROW_MAX);
      r2 = (long long)ceilf(g * (float)Ro);
    }
    Rn = r2 > Rcover ? Rcover : (int)r2;
  }
  if (Rn > Rcover) Rn = Rcover;
  if (Rn <= Ro) Rn = Ro + 1;
  return Rn;
}

__device__ __forceinline__ int cover_radius(const CloudInfo &ci, int4 qc) {
  int r = max(qc.x, ci.nx - 1 - qc.x);
  r = max(r, max(qc.y, ci.ny - 1 - qc.y));
  r = max(r, max(qc.z, ci.nz - 1 - qc.z));
  return r < 0 ? 0 : r;
}

__device__ __forceinline__ uint32_t pad_value(const QueryArgs &a, float4 qr,
                                              int qi) {
  if (a.pad_mode == FK_PAD_SELF) return __float_as_uint(qr.w);
  if (a.pad_mode == FK_PAD_MINUS1) return 0xffffffffu;
  return (uint32_t)a.qpad[qi];
}

// Output slot of row `row` (batched: per-cloud K and row base).
__device__ __forceinline__ int64_t out_base(const QueryArgs &a,
                                            const CloudInfo &ci, int cloud,
                                            int row, int k) {
  if (a.k_per_cloud)
    return a.out_offsets[cloud] + (int64_t)(row - ci.first_id) * k;
  return (int64_t)row * k;
}

// One output entry (index + distance) from a key.
__device__ __forceinline__ void write_entry(const QueryArgs &a, int64_t o,
                                            uint64_t key) {
  const int id = (int)key_id(key);
  if (a.idx64)
    ((int64_t *)a.out_idx)[o] = (int64_t)id;
  else
    ((int32_t *)a.out_idx)[o] = id;
  if (a.out_d) {
    const uint32_t db = key_dbits(key);
    float d = __uint_as_float(db);
    if (db == kInfBits)
      d = a.pad_dist;
    else if (a.dist_mode == FK_DIST_SQRT)
      d = sqrtf(d);
    a.out_d[o] = d;
  }
}

template <int KC, int G>
__device__ __forceinline__ void write_row(const QueryArgs &a, const TopK<KC> &t,
                                          int64_t o, int sub) {
  // real entry j lives in register slot base + j; lane `sub` of the group
  // writes the entries j with j % G == sub
  const int64_t ob = o - t.base();
#pragma unroll
  for (int j = 0; j < KC; ++j)
    if (j >= t.base() && ((j - t.base()) & (G - 1)) == sub) write_entry(a, ob + j, t.top[j]);
}

__device__ __forceinline__ int wave_prefix_slot(bool pred, int32_t *counter) {
  const uint64_t m = __ballot(pred);
  const int lane = threadIdx.x & (kWave - 1);
  int base = 0;
  if (m != 0ull) {
    const int leader = __builtin_ctzll(m);
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader, kWave);
  }
  const uint64_t below = lane ? (m & (~0ull >> (64 - lane))) : 0ull;
  return base + __popcll(below);
}

// An uncertified query the first pass finishes itself instead of listing it
// for the expansion launch: one per tile at most, only when it is the tile's
// only one and already holds k candidates (its next shell is sized from its
// k-th distance), and at most kInlineBudget per launch.  That one query's
// shells are a chain of dependent reads (~7 us alone in the expansion
// launch) which a first-pass wave hides behind the other waves; many such
// queries, or ones still searching for k candidates, would lengthen the
// first pass instead.

template <int KC>
__device__ __forceinline__ uint64_t *inline_seed(uint64_t *wave_buf, uint64_t *wave_tab);
template <int KC>
__device__ __forceinline__ void inline_expand(const QueryArgs &a, int q, int lane,
                                              uint64_t *wave_buf, uint64_t *wave_tab);

// One wave: queries [q0, q0 + 64 / G) of the sorted query list, G lanes per
// query, all of one K class.  Used by the plain kernel and by the persistent
// batched kernel.  With INL (the plain kernel) it returns the sorted index of
// a query left for the wave to expand (its list in the wave's seed area), or
// -1.
// First pending-list slot of the launch's list (per K class in batched
// mode, written by k_items); the first-pass kernels take lists already
// offset (pend_class -1).
__device__ __forceinline__ int pend_base(const QueryArgs &a) {
  return a.pend_class >= 0 ? a.scalars[kPendBase + a.pend_class] : 0;
}

template <int KC, int G, bool INL = false, int CB = default_batch<KC>()>
__device__ __forceinline__ int run_tile(const QueryArgs &a, int q0, int qend,
                                        uint64_t *lane_buf, uint64_t *lane_tab) {
  const int lane = threadIdx.x & (kWave - 1);
  const int sub = lane & (G - 1);
  const int qi = q0 + lane / G;
  const bool active = qi < qend;
  const int qs = active ? qi : (qend - 1);
  const float4 qr = a.qrec[qs];
  const int4 qc = a.qcell[qs];
  const CloudInfo ci = a.clouds[qc.w];
  const int k = a.k_per_cloud ? a.k_per_cloud[qc.w] : a.k;
  TopK<KC> t;
  t.init(lane_buf, k, a.cut_r2, pad_value(a, qr, qs));
  scan_block<KC, G, CB>(t, a, ci, qc, qr, a.R, active, sub, lane, lane_tab);
  if (wave_any(t.cnt > 0)) t.flush();
  if constexpr (G > 1) t.template group_merge<G>(lane);

  bool done = true;
  if (a.exact) {
    double umax;
    const double gap = cert_gap(ci, qc, qr, a.R, umax);
    float eff = __uint_as_float(key_dbits(t.kth()));
    eff = eff < a.cut_r2 ? eff : a.cut_r2;
    done = certified(gap, eff, ci.inv_cell, umax, a.R);
    // one pending slot per uncertified query: its group's lane 0 (the
    // certified count is the query count minus the pending one, so certified
    // queries touch no shared counter)
    const bool lead = sub == 0;
    if constexpr (INL && kInlineBudget > 0) {
      const uint64_t mu = __ballot(active && !done && lead);
      const bool found = key_dbits(t.kth()) != kInfBits;
      if (__popcll(mu) == 1 && __ballot(active && !done && lead && found) == mu) {
        int claim = 0;
        if (lane == 0) claim = atomicAdd(a.qcnt + kCntInline, 1);
        claim = __builtin_amdgcn_readfirstlane(claim);
        if (claim < kInlineBudget) {  // wave-uniform
          const int lead_lane = __builtin_ctzll(mu);
          uint64_t *seed = inline_seed<KC>(lane_buf - lane, lane_tab - lane);
          wave_lds_sync();  // the tile's last merge has read its buffers
          if (lane == lead_lane) {
#pragma unroll
            for (int j = 0; j < KC; ++j)
              if (j >= t.base()) seed[j - t.base()] = t.top[j];
          }
          if (active && done)
            write_row<KC, G>(a, t, out_base(a, ci, qc.w, __float_as_int(qr.w), k), sub);
          return __builtin_amdgcn_readlane(qs, lead_lane);
        }
      }
    }
    int slot = wave_prefix_slot(active && !done && lead, a.qcnt + kCntPending);
    if constexpr (G > 1) slot = __shfl(slot, lane & ~(G - 1), kWave);
    if (active && !done && slot >= a.nq) {
      // the pending counter was not reset before this pass: refuse to write
      if (lead) raise_status(a.scalars, FK_ERR_STATE);
    } else if (active && !done) {
      if (lead) a.pend_q[slot] = qs;
      const int64_t pb = (int64_t)slot * a.k - t.base();
#pragma unroll
      for (int j = 0; j < KC; ++j)
        if (j >= t.base() && ((j - t.base()) & (G - 1)) == sub)
          a.pend_keys[pb + j] = t.top[j];
    }
  }
  if (active && done) {
    const int row = __float_as_int(qr.w);
    write_row<KC, G>(a, t, out_base(a, ci, qc.w, row, k), sub);
  }
  return -1;
}

template <int KC>
constexpr int waves_per_block() {
  return KC <= 16 ? 4 : (KC == 32 ? 2 : 1);
}

// G lanes per query (the default per K class, or 1 with FK_QUERY_ONE_LANE),
// CB records per lane per scan round.
template <int KC, int G, int CB = default_batch<KC>()>
__global__ void __launch_bounds__(64 * waves_per_block<KC>(), min_waves_per_eu<KC>())
k_query(QueryArgs a) {
  constexpr int WPB = waves_per_block<KC>();
  constexpr int BS = buffer_slots<KC>();
  __shared__ uint64_t sbuf[WPB * BS * kWave];
  __shared__ uint64_t stab[WPB * stream_entries<KC>() * kWave];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  if (blockIdx.x == 0 && threadIdx.x < kCntStride && a.qcnt_next) a.qcnt_next[threadIdx.x] = 0;
  const int q0 = (blk * WPB + w) * (kWave / G);
  // wave-uniform exits: past the end, or the build/attach raised an error
  // (its tables may then be incomplete)
  if (q0 >= a.nq || a.scalars[kStatus] != 0) return;
  uint64_t *wb = sbuf + w * BS * kWave, *wt = stab + w * stream_entries<KC>() * kWave;
  const int qx = run_tile<KC, G, (FK_INLINE_PLAIN != 0), CB>(a, q0, a.nq, wb + lane, wt + lane);
  if constexpr (FK_INLINE_PLAIN != 0)
    if (qx >= 0) inline_expand<KC>(a, qx, lane, wb, wt);
}

// The balanced first pass: wave w runs plan[w] = (first query, G) with G
// lanes per query (see launch_plan).  One kernel serves every G, so a
// launch mixes wide groups for the dense super-tiles with narrow ones for
// the sparse.
template <int KC, int CB = default_batch<KC>()>
__global__ void __launch_bounds__(64 * waves_per_block<KC>(), min_waves_per_eu<KC>())
k_query_planned(QueryArgs a) {
  constexpr int WPB = waves_per_block<KC>();
  constexpr int BS = buffer_slots<KC>();
  __shared__ uint64_t sbuf[WPB * BS * kWave];
  __shared__ uint64_t stab[WPB * stream_entries<KC>() * kWave];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & (kWave - 1);
  if (blockIdx.x == 0 && threadIdx.x < kCntStride && a.qcnt_next) a.qcnt_next[threadIdx.x] = 0;
  const int wid = blockIdx.x * WPB + w;
  if (wid >= a.plan_waves || a.scalars[kStatus] != 0) return;
  const int2 d = a.plan[wid];
  const int q0 = __builtin_amdgcn_readfirstlane(d.x);
  const int g = __builtin_amdgcn_readfirstlane(d.y);
  if (q0 >= a.nq) return;
  uint64_t *wb = sbuf + w * BS * kWave, *wt = stab + w * stream_entries<KC>() * kWave;
  const int qend = min(a.nq, q0 + kWave / g);
  int qx;
  switch (g) {
    case 1: qx = run_tile<KC, 1, true, CB>(a, q0, qend, wb + lane, wt + lane); break;
    case 2: qx = run_tile<KC, 2, true, CB>(a, q0, qend, wb + lane, wt + lane); break;
    case 4: qx = run_tile<KC, 4, true, CB>(a, q0, qend, wb + lane, wt + lane); break;
    default: qx = run_tile<KC, 8, true, CB>(a, q0, qend, wb + lane, wt + lane); break;
  }
  if (qx >= 0) inline_expand<KC>(a, qx, lane, wb, wt);
}

// ------------------------------------------- adaptive neighbourhood query
//
// Point p of the Morton-sorted order scans the 27 cells around its cell at
// its selected level (FlashKNN/functions/adaptive_octree.py:369-488 flattens
// the same query into one launch).  A group of G lanes serves one point and
// one wave 64 / G consecutive points, which mostly share cells and levels.
// Lane `sub` of the group looks up the cells j = sub, sub + G, ... of the 27
// in its level's table (the lookups in flight together); the runs then reach
// every lane of the group by shuffles, nine cells at a time, and lane `sub`
// takes every G-th record of them (the group's loads adjacent, its lanes'
// work even).  (Two lanes each looking up all 27 cells: 160 us for 92k
// points; four lanes each streaming whole runs of its own cells: 214 us.)
// Rows go to the input order; slots without a candidate hold (the point
// itself, +inf).
#ifndef FK_AD_G
#define FK_AD_G 4
#endif
constexpr int kAdG = FK_AD_G;
constexpr int kAdCells = (27 + kAdG - 1) / kAdG;  // cells looked up per lane

template <int KC>
__global__ void __launch_bounds__(64 * waves_per_block<KC>(), min_waves_per_eu<KC>())
k_adaptive_query(AdaptiveArgs a) {
  constexpr int WPB = waves_per_block<KC>();
  constexpr int BS = buffer_slots<KC>();
  constexpr int G = kAdG;
  constexpr int NC = kAdCells;
  constexpr int S = kSlots <= BS ? kSlots : BS;
  static_assert(G >= 2 && G <= 16 && (G & (G - 1)) == 0, "adaptive query: lanes per point");
  __shared__ uint64_t sbuf[WPB * BS * kWave];
  __shared__ uint64_t stab[WPB * kStreamCols * kWave];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
  const int sub = lane & (G - 1), gbase = lane & ~(G - 1);
  const int q0 = (xcd_remap(blockIdx.x, gridDim.x) * WPB + w) * (kWave / G);
  // wave-uniform exits: past the end, or the selection raised an error
  if (q0 >= a.n || a.stats[kAdStatus] != 0) return;
  const int qi = q0 + lane / G;
  const bool active = qi < a.n;
  const int qs = active ? qi : a.n - 1;
  const float4 qr = a.srec[qs];
  const int s = a.level[qs];
  const uint64_t c = a.skeys[qs] >> (3 * s);
  AdNeighbours nb;
  nb.init(c, s);
  uint64_t cells[NC];
#pragma unroll
  for (int i = 0; i < NC; ++i) cells[i] = kAdEmpty;
#pragma unroll
  for (int j = 0; j < 27; ++j)
    if (active && (j & (G - 1)) == sub) cells[j / G] = nb.key(j);
  AdEntry own[NC];
  int64_t slot[NC];
  ad_find_n<NC>(a.ent, a.lg, s, cells, own, slot);
  uint64_t *lane_buf = sbuf + w * BS * kWave + lane;
  uint64_t *lane_tab = stab + w * kStreamCols * kWave + lane;
  TopK<KC> t;
  t.init(lane_buf, a.k, a.cut_r2, __float_as_uint(qr.w));
#pragma unroll
  for (int part = 0; part < 3; ++part) {
    int b[9], e[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      const int j = 9 * part + m;
      const int from = gbase + (j & (G - 1));
      b[m] = __shfl(own[j / G].start, from, kWave);
      e[m] = b[m] + __shfl(own[j / G].count, from, kWave);
    }
    ColumnStream<9, G, S> cs;
    cs.init(lane_tab, b, e, sub);
    scan_stream<KC, G, 9, S>(t, cs, qr, a.srec);
  }
  if (wave_any(t.cnt > 0)) t.flush();
  t.template group_merge<G>(lane);
  if (!active) return;
  const int64_t ob = (int64_t)__float_as_int(qr.w) * a.k - t.base();
#pragma unroll
  for (int j = 0; j < KC; ++j) {
    if (j >= t.base() && ((j - t.base()) & (G - 1)) == sub) {
      const uint64_t key = t.top[j];
      a.out_idx[ob + j] = (int64_t)key_id(key);
      if (a.out_d) a.out_d[ob + j] = __uint_as_float(key_dbits(key));
    }
  }
}

// ------------------------------------------------- cooperative expansion
//
// Pending queries of FK_MODE_EXACT (not certified by the first pass): one
// wave per query.  The query's best-k list is spread over the wave, lane i
// holding the i-th best key; the 64 lanes split the shell's columns between
// them, append the candidates below the current k-th key to a per-wave LDS
// buffer, and every 64 buffered keys are merged by a wave-wide bitonic sort
// and merge through lane shuffles.  The block then grows shell by shell
// (radius from the current k-th distance, doubled while it is infinite)
// until the query certifies or the block covers its cloud's box.

constexpr int kCoopWaves = 4;          // waves per block
constexpr int kCoopCap = 256;          // buffered keys per wave


// Ascending bitonic sort of one key per lane across the wave.
__device__ __forceinline__ uint64_t wave_sort(uint64_t v, int lane) {
#pragma unroll
  for (int k2 = 2; k2 <= kWave; k2 <<= 1) {
#pragma unroll
    for (int j = k2 >> 1; j > 0; j >>= 1) {
      const uint64_t o = lane_xor_n(v, j);
      const bool up = (lane & k2) == 0;
      const bool lower = (lane & j) == 0;
      v = (lower == up) ? kmin(v, o) : kmax(v, o);
    }
  }
  return v;
}

// L and S ascending across lanes -> the 64 smallest of both, ascending.
__device__ __forceinline__ uint64_t wave_merge(uint64_t L, uint64_t S, int lane) {
  uint64_t v = kmin(L, lane_reverse(S));
#pragma unroll
  for (int j = kWave >> 1; j > 0; j >>= 1) {
    const uint64_t o = lane_xor_n(v, j);
    v = (lane & j) == 0 ? kmin(v, o) : kmax(v, o);
  }
  return v;
}

// Bitonic sequence across the wave -> ascending (butterfly of 6 stages).
__device__ __forceinline__ uint64_t wave_bitonic_merge(uint64_t v, int lane) {
#pragma unroll
  for (int j = kWave >> 1; j > 0; j >>= 1) {
    const uint64_t o = lane_xor_n(v, j);
    v = (lane & j) == 0 ? kmin(v, o) : kmax(v, o);
  }
  return v;
}

// Best list of P * 64 keys: L[p] at lane i holds rank p * 64 + i.
template <int P>
struct CoopList {
  uint64_t L[P];
  uint64_t *buf;    // the wave's LDS buffer
  int cnt;          // buffered keys (wave-uniform)
  int k;
  uint64_t cutkey, thr;

  // k is wave-uniform: the k-th key is read from its lane with v_readlane
  __device__ __forceinline__ uint64_t kth() const {
    const int r = __builtin_amdgcn_readfirstlane(k - 1);
    uint64_t v = L[0];
    int src = r;
    if constexpr (P > 1) {
      if (r >= kWave) {
        v = L[1];
        src = r - kWave;
      }
    }
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, src);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
  }
  __device__ __forceinline__ void refresh(int lane) {
    (void)lane;
    const uint64_t kk = kth();
    thr = kk < cutkey ? kk : cutkey;
  }
  __device__ __forceinline__ void flush(int lane) {
    for (int c0 = 0; c0 < cnt; c0 += kWave) {
      const uint64_t v = c0 + lane < cnt ? buf[c0 + lane] : kMaxKey;
      const uint64_t S = wave_sort(v, lane);
      if constexpr (P == 1) {
        L[0] = wave_merge(L[0], S, lane);
      } else {
        // 128 smallest of L (128, ascending) and S (64, ascending; padded
        // with +inf to 128): C[i] = min(L[i], S'[127 - i]) is bitonic with
        // C[0..63] = L[0..63]; the first merge stage (distance 64) is the
        // in-lane min/max, then each half merges across the wave.
        const uint64_t c1 = kmin(L[1], lane_reverse(S));
        const uint64_t lo = kmin(L[0], c1), hi = kmax(L[0], c1);
        L[0] = wave_bitonic_merge(lo, lane);
        L[1] = wave_bitonic_merge(hi, lane);
      }
    }
    cnt = 0;
    refresh(lane);
  }
  __device__ __forceinline__ void offer(uint64_t key, bool has, int lane) {
    const bool acc = has && key < thr;
    const uint64_t m = __ballot(acc);
    if (m != 0ull) {
      const uint64_t below = lane ? (m & (~0ull >> (64 - lane))) : 0ull;
      if (acc) buf[cnt + __popcll(below)] = key;
      cnt += __popcll(m);
      if (cnt > kCoopCap - kWave) flush(lane);
    }
  }
};

// Columns of the shell Ro < Chebyshev distance <= Rn (Ro = -1: the whole
// block of half-width Rn).  Each lane looks up kShellBatch of the shell's
// columns at a time (their cell-table reads go out together); the records
// of the batch's non-empty column segments (an inner column has two) are
// then spread evenly over the wave: the segments go to the wave's LDS table
// as (record prefix, first record), lane l takes records [l C, (l + 1) C)
// of their concatenation, C = ceil(total / 64), found by one binary search
// and then walked in order.  (With one lane per column the wave waited for
// its longest column: a room query's 15-column shell took 8 dependent
// rounds of four loads, LiDAR shells through tall objects up to 100.)
// Four columns per lane at a time: with two expansion waves per SIMD, 4
// beat 6, 8, 12 and 16 (the smaller table also leaves the batched class
// kernels more LDS beside the expansion).
#ifndef FK_SHELL_BATCH
#define FK_SHELL_BATCH 4
#endif
constexpr int kShellBatch = FK_SHELL_BATCH;  // column segments looked up together per lane
constexpr int kShellTab = 2 * kShellBatch * kWave;  // segment table entries per wave

// Exclusive prefix sum over the wave (every lane active); *total = the sum.
__device__ __forceinline__ int wave_excl_scan(int v, int lane, int *total) {
  int x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int u = __shfl_up(x, o, kWave);
    if (lane >= o) x += u;
  }
  *total = __builtin_amdgcn_readlane(x, kWave - 1);
  return x - v;
}

template <int P>
__device__ __forceinline__ void coop_shell(CoopList<P> &t, const QueryArgs &a, const CloudInfo &ci,
                           int4 qc, float4 qr, int Ro, int Rn, int lane, uint64_t *wtab) {
  const int x0 = max(-Rn, -qc.x), x1 = min(Rn, ci.nx - 1 - qc.x);
  const int y0 = max(-Rn, -qc.y), y1 = min(Rn, ci.ny - 1 - qc.y);
  const long long wy = y1 >= y0 ? (long long)(y1 - y0 + 1) : 0;
  const long long ncol = x1 >= x0 ? (long long)(x1 - x0 + 1) * wy : 0;
  auto zc = [&](long long v) {
    return (int)(v < -1 ? -1 : (v > ci.nz ? ci.nz : v));
  };
  const long long zq = qc.z;
  const int zlo = zc(zq - Rn), zhi = zc(zq + Rn);
  const int zin_lo = zc(zq - Ro - 1), zin_hi = zc(zq + Ro + 1);
  long long s = lane;  // this lane's next column
  // (row, offset in row) of column s, stepped by kWave columns at a time
  // without dividing (64-bit divisions per column dominated sparse shells)
  const int wyi = wy > 0 ? (int)(wy < INT_MAX ? wy : INT_MAX) : 1;
  const int step_q = kWave / wyi, step_r = kWave % wyi;
  int cq = lane / wyi, cr = lane % wyi;
  while (wave_any(s < ncol)) {
    int b[2 * kShellBatch], en[2 * kShellBatch];
#pragma unroll
    for (int j = 0; j < kShellBatch; ++j) {
      const long long sj = s + (long long)j * kWave;
      b[2 * j] = en[2 * j] = b[2 * j + 1] = en[2 * j + 1] = 0;
      const int dx = x0 + cq, dy = y0 + cr;  // column sj
      cr += step_r;
      cq += step_q;
      if (cr >= wyi) {
        cr -= wyi;
        ++cq;
      }
      if (sj < ncol) {
        const bool inner = dx >= -Ro && dx <= Ro && dy >= -Ro && dy <= Ro;
        column_range(ci, a.cell_start, a.scell, qc.x + dx, qc.y + dy, zlo, inner ? zin_lo : zhi,
                     b[2 * j], en[2 * j]);
        if (inner)
          column_range(ci, a.cell_start, a.scell, qc.x + dx, qc.y + dy, zin_hi, zhi,
                       b[2 * j + 1], en[2 * j + 1]);
      }
    }
    s += (long long)kShellBatch * kWave;
    int qn = 0, tot = 0;
#pragma unroll
    for (int j = 0; j < 2 * kShellBatch; ++j)
      if (b[j] < en[j]) {
        ++qn;
        tot += en[j] - b[j];
      }
    int nr, T;
    int q = wave_excl_scan(qn, lane, &nr);
    int pre = wave_excl_scan(tot, lane, &T);
#pragma unroll
    for (int j = 0; j < 2 * kShellBatch; ++j)
      if (b[j] < en[j]) {
        put_range(wtab + q, pre, b[j]);  // (first record's place, first record)
        pre += en[j] - b[j];
        ++q;
      }
    wave_lds_sync();
    if (T > 0) {  // wave-uniform
      const int C = (T + kWave - 1) / kWave;
      int r = lane * C;
      const int rend = min(r + C, T);
      // the segment holding record r: the last i with prefix <= r
      int lo = 0, hi = nr - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (get_range(wtab + mid).x <= r) lo = mid; else hi = mid - 1;
      }
      int i = lo;
      int2 cur = get_range(wtab + i);
      int nxt = i + 1 < nr ? get_range(wtab + i + 1).x : T;
      // four records of the lane's stretch in flight per step
      while (wave_any(r < rend)) {
        float4 rec[4];
        bool has[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          has[u] = r + u < rend;
          while (has[u] && r + u >= nxt) {
            ++i;
            cur = get_range(wtab + i);
            nxt = i + 1 < nr ? get_range(wtab + i + 1).x : T;
          }
          rec[u] = a.srec[has[u] ? (uint32_t)(cur.y + (r + u - cur.x)) : 0u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint64_t key =
              has[u] ? make_key(sqdist(rec[u].x, rec[u].y, rec[u].z, qr.x, qr.y, qr.z),
                                __float_as_uint(rec[u].w))
                     : kMaxKey;
          t.offer(key, has[u], lane);
        }
        r += 4;
      }
    }
    wave_lds_sync();  // the next batch rewrites the table
  }
  if (t.cnt > 0) t.flush(lane);
}

// Real keys in a cooperative list (wave-uniform).
template <int P>
__device__ __forceinline__ int coop_found(const CoopList<P> &t, int lane) {
  int n = 0;
#pragma unroll
  for (int p = 0; p < P; ++p)
    n += __popcll(__ballot(p * kWave + lane < t.k && key_dbits(t.L[p]) != kInfBits));
  return n;
}

// min(k-th distance, cut) of a cooperative list.
template <int P>
__device__ __forceinline__ float coop_eff(const CoopList<P> &t, float cut_r2) {
  const float e = __uint_as_float(key_dbits(t.kth()));
  return e < cut_r2 ? e : cut_r2;
}

// Sum over the wave (every lane active), by DPP and row swaps.
__device__ __forceinline__ int wave_sum(int v) {
  v += (int)lane_xor32<1>((uint32_t)v);
  v += (int)lane_xor32<2>((uint32_t)v);
  v += (int)lane_xor32<4>((uint32_t)v);
  v += (int)lane_xor32<8>((uint32_t)v);
  v += (int)lane_xor32<16>((uint32_t)v);
  v += (int)lane_xor32<32>((uint32_t)v);
  return v;
}

// Half-widths the radius probe tries (at most 31: one lane per row), in
// two rounds of loads (the second only when the first finds nothing).
constexpr int kProbeRadii[] = {2, 3, 4, 6, 8, 11, 16, 22, 31};
constexpr int kNumProbes = sizeof(kProbeRadii) / sizeof(kProbeRadii[0]);
constexpr int kProbeSplit = 4;

// Probes J0 .. J1 - 1 (see probe_radius); 0 if none qualifies.
template <int J0, int J1>
__device__ __forceinline__ int probe_round(const CloudInfo &ci, const int32_t *__restrict__ cs,
                                           int4 qc, int Ro, int Rcover, int k, int dx) {
  const int x = qc.x + dx;
  const bool row = x >= 0 && x < ci.nx;
  int cnt[J1 - J0];
#pragma unroll
  for (int j = J0; j < J1; ++j) {
    const int R = kProbeRadii[j] < Rcover ? kProbeRadii[j] : Rcover;
    cnt[j - J0] = 0;
    if (R > Ro && row && dx >= -R && dx <= R) {
      const int y0 = max(qc.y - R, 0), y1 = min(qc.y + R, ci.ny - 1);
      if (y0 <= y1) cnt[j - J0] = cs[lin_of(ci, x, y1 + 1, 0)] - cs[lin_of(ci, x, y0, 0)];
    }
  }
  int best = 0;
#pragma unroll
  for (int j = J1 - 1; j >= J0; --j) {
    const int R = kProbeRadii[j] < Rcover ? kProbeRadii[j] : Rcover;
    if (R > Ro && wave_sum(cnt[j - J0]) >= k) best = R;
  }
  return best;
}

// Wave-uniform.  For a query that holds fewer than k candidates (no k-th
// distance to size the next block): the smallest probe half-width R above
// Ro (clipped to Rcover) whose block holds at least k points when counted
// over the full height of the box, or 0 if none does.  The count of
// columns (x, y0..y1) over all z is the difference of two cell-table
// entries (a row's cells are consecutive, z fastest; the column layout has
// one entry per column), so lane l sums row x = qx + l - 32 for several
// probes at once: one or two rounds of loads instead of the shell-by-shell
// growth through empty space around an isolated query (a far LiDAR return
// needed up to nine shells to reach its neighbours 8 m away).  The count
// is an upper bound of the block's, so the block may still hold fewer
// than k; the growth then goes on from it.
__device__ __forceinline__ int probe_radius(const CloudInfo &ci, const int32_t *__restrict__ cs,
                                            int4 qc, int Ro, int Rcover, int k, int lane) {
  const int dx = lane - 32;
  const int r = probe_round<0, kProbeSplit>(ci, cs, qc, Ro, Rcover, k, dx);
  return r ? r : probe_round<kProbeSplit, kNumProbes>(ci, cs, qc, Ro, Rcover, k, dx);
}

// The next half-width of a cooperative list's block (wave-uniform): from
// the k-th distance (or the cut) when there is one, else from the radius
// probe, else by the growth rule of next_radius.  The results never depend
// on it: every shell sequence ends with the same certified block contents.
template <int P>
__device__ __forceinline__ int coop_next_radius(const CoopList<P> &t, const QueryArgs &a,
                                                const CloudInfo &ci, int4 qc, float4 qr,
                                                int Ro, int Rcover, int lane) {
  const float eff = coop_eff(t, a.cut_r2);
  if (!(eff < INFINITY)) {
    const int Rp = probe_radius(ci, a.cell_start, qc, Ro, Rcover, t.k, lane);
    if (Rp > Ro) return Rp;
  }
  return next_radius(ci, qc, qr, Ro, eff, Rcover, coop_found(t, lane), t.k);
}

// Grow the block from half-width Ro shell by shell until the query
// certifies or the block covers the cloud's box.  Returns the final Ro.
template <int P>
__device__ __forceinline__ int coop_expand(CoopList<P> &t, const QueryArgs &a,
                                           const CloudInfo &ci, int4 qc, float4 qr,
                                           int Ro, int lane, uint64_t *stab) {
  const int Rcover = cover_radius(ci, qc);
  for (;;) {  // wave-uniform
    double umax;
    const double gap = cert_gap(ci, qc, qr, Ro, umax);
    if (certified(gap, coop_eff(t, a.cut_r2), ci.inv_cell, umax, Ro) || Ro >= Rcover)
      return Ro;
    const int Rn = next_radius(ci, qc, qr, Ro, coop_eff(t, a.cut_r2), Rcover,
                               coop_found(t, lane), t.k);
    coop_shell<P>(t, a, ci, qc, qr, Ro, Rn, lane, stab);
    Ro = Rn;
  }
}

#if FK_EXPAND_PART
__global__ void __launch_bounds__(64 * kCoopWaves, 4) k_expand_coop(QueryArgs a) {
  __shared__ uint64_t sbuf[kCoopWaves * kCoopCap];
  __shared__ uint64_t ctab[kCoopWaves * kShellTab];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
  uint64_t *wave_tab = ctab + w * kShellTab;
  const int base = pend_base(a);
  const int npend = a.scalars[kStatus] != 0 ? 0 : min(a.qcnt[kCntPending], a.nq - base);
  int nexp = 0;
  // static striding over the pending list (a shared work counter measured
  // slower: its atomics cost more than the balance they buy)
  const int nwaves = gridDim.x * kCoopWaves;
  for (int i = blockIdx.x * kCoopWaves + w; i < npend; i += nwaves) {
    const int slot = base + i;
    const int qs = a.pend_q[slot];
    const float4 qr = a.qrec[qs];
    const int4 qc = a.qcell[qs];
    const CloudInfo ci = a.clouds[qc.w];
    const int k = a.k_per_cloud ? a.k_per_cloud[qc.w] : a.k;
    CoopList<1> t;
    t.buf = sbuf + w * kCoopCap;
    t.cnt = 0;
    t.k = k;
    t.cutkey = make_key(a.cut_r2, 0u);
    t.L[0] = lane < k ? a.pend_keys[(int64_t)slot * a.k + lane] : kMaxKey;
    t.refresh(lane);
    // the first pass left it uncertified at half-width a.R; the first
    // shell is sized by the radius probe when the list is short of k (once:
    // inside the growth loop the probe's registers spilled the kernel)
    const int Rcover = cover_radius(ci, qc);
    if (a.R < Rcover) {
      const int Rn = coop_next_radius<1>(t, a, ci, qc, qr, a.R, Rcover, lane);
      coop_shell<1>(t, a, ci, qc, qr, a.R, Rn, lane, wave_tab);
      coop_expand<1>(t, a, ci, qc, qr, Rn, lane, wave_tab);
    }
    ++nexp;
    if (lane < k)
      write_entry(a, out_base(a, ci, qc.w, __float_as_int(qr.w), k) + lane, t.L[0]);
  }
  if (lane == 0 && nexp) atomicAdd(a.qcnt + kCntExpanded, nexp);
}
#endif  // FK_COMMON

// The tile's LDS (key buffers, stream table) once its scan is done: the
// shell table at the start of the key buffers, then the list buffer and the
// seed (the inline query's first-pass list), in the key buffers when they
// are long enough, else in the stream table.
template <int KC>
struct InlineLds {
  static constexpr int kBuf = buffer_slots<KC>() * kWave;
  static constexpr int kTab = stream_entries<KC>() * kWave;
  static constexpr bool kInBuf = kBuf >= kShellTab + kCoopCap + kWave;
  static_assert(kBuf >= kShellTab, "tile buffers below the shell table");
  static_assert(kInBuf || kTab >= kCoopCap + kWave, "tile LDS below the list buffer");
  static __device__ __forceinline__ uint64_t *coop(uint64_t *wave_buf, uint64_t *wave_tab) {
    return kInBuf ? wave_buf + kShellTab : wave_tab;
  }
};

template <int KC>
__device__ __forceinline__ uint64_t *inline_seed(uint64_t *wave_buf, uint64_t *wave_tab) {
  return InlineLds<KC>::coop(wave_buf, wave_tab) + kCoopCap;
}

// The query run_tile left to its wave (sorted index q, list in the seed
// area), finished as k_expand_coop finishes a listed one: the shells grow
// until it certifies, then its row is written.
template <int KC>
__device__ __forceinline__ void inline_expand(const QueryArgs &a, int q, int lane,
                                              uint64_t *wave_buf, uint64_t *wave_tab) {
  uint64_t *coop = InlineLds<KC>::coop(wave_buf, wave_tab);
  uint64_t *seed = coop + kCoopCap;
  wave_lds_sync();  // the seed is written
  const float4 qr = a.qrec[q];
  const int4 qc = a.qcell[q];
  const CloudInfo ci = a.clouds[qc.w];
  const int k = a.k_per_cloud ? a.k_per_cloud[qc.w] : a.k;
  CoopList<1> c;
  c.buf = coop;
  c.cnt = 0;
  c.k = k;
  c.cutkey = make_key(a.cut_r2, 0u);
  c.L[0] = lane < k ? seed[lane] : kMaxKey;
  c.refresh(lane);
  const int Rcover = cover_radius(ci, qc);
  if (a.R < Rcover) {
    const int Rn = next_radius(ci, qc, qr, a.R, coop_eff(c, a.cut_r2), Rcover,
                               coop_found(c, lane), k);
    coop_shell<1>(c, a, ci, qc, qr, a.R, Rn, lane, wave_buf);
    coop_expand<1>(c, a, ci, qc, qr, Rn, lane, wave_buf);
  }
  if (lane < k)
    write_entry(a, out_base(a, ci, qc.w, __float_as_int(qr.w), k) + lane, c.L[0]);
  if (lane == 0) atomicAdd(a.qcnt + kCntExpanded, 1);
}

// k in (64, 128]: one wave per query from the start, the best list held as
// two keys per lane.  Scans the block of half-width a.R; FK_MODE_EXACT then
// grows it until the query certifies.
#if FK_WIDE_PART
__global__ void __launch_bounds__(64 * kCoopWaves) k_query_wide(QueryArgs a) {
  __shared__ uint64_t sbuf[kCoopWaves * kCoopCap];
  __shared__ uint64_t ctab[kCoopWaves * kShellTab];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
  uint64_t *wave_tab = ctab + w * kShellTab;
  if (blockIdx.x == 0 && threadIdx.x < kCntStride && a.qcnt_next) a.qcnt_next[threadIdx.x] = 0;
  const int nq = a.scalars[kStatus] != 0 ? 0 : a.nq;
  const int nwaves = gridDim.x * kCoopWaves;
  int ncert = 0, nexp = 0;
  for (int qs = blockIdx.x * kCoopWaves + w; qs < nq; qs += nwaves) {
    const float4 qr = a.qrec[qs];
    const int4 qc = a.qcell[qs];
    const CloudInfo ci = a.clouds[qc.w];
    const int k = a.k;
    CoopList<2> t;
    t.buf = sbuf + w * kCoopCap;
    t.cnt = 0;
    t.k = k;
    t.cutkey = make_key(a.cut_r2, 0u);
    const uint64_t padkey = key_of_bits(kInfBits, pad_value(a, qr, qs));
    t.L[0] = lane < k ? padkey : kMaxKey;
    t.L[1] = lane + kWave < k ? padkey : kMaxKey;
    t.refresh(lane);
    coop_shell<2>(t, a, ci, qc, qr, -1, a.R, lane, wave_tab);
    if (a.exact) {
      const int Ro = coop_expand<2>(t, a, ci, qc, qr, a.R, lane, wave_tab);
      if (Ro > a.R) ++nexp; else ++ncert;
    }
    const int64_t o = out_base(a, ci, qc.w, __float_as_int(qr.w), k);
    if (lane < k) write_entry(a, o + lane, t.L[0]);
    if (lane + kWave < k) write_entry(a, o + kWave + lane, t.L[1]);
  }
  // expanded queries count as pending too: certified = queries - pending
  (void)ncert;
  if (lane == 0 && nexp) {
    atomicAdd(a.qcnt + kCntPending, nexp);
    atomicAdd(a.qcnt + kCntExpanded, nexp);
  }
}
#endif  // FK_COMMON

// ----------------------------------------------------- batched persistent

// Work item i -> (cloud, first query of the tile) using the per-cloud tile
// prefix item_cloud[c] (exclusive, int32[num_clouds + 1]).
__device__ __forceinline__ int item_to_cloud(const int32_t *prefix, int B,
                                             int item) {
  int lo = 0, hi = B;  // largest c with prefix[c] <= item
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}

// The batched launch's K classes (k up to 8, 16, 32, 64 -> 0..3) and
// their work items (kWave / G queries of one cloud).  The batched kernels
// take one lane per query (FK_BATCHED_G): a batch holds many more tiles
// than wave slots, so per-query cost, not a wave's latency, sets the step,
// and one lane per query is the cheapest per query (16-cloud batch: 0.519
// ms with two lanes for K classes 8/16, 0.490 ms with one; K class 32
// alone 0.509).
#ifndef FK_BATCHED_G
#define FK_BATCHED_G 1
#endif
template <int KC>
__host__ __device__ constexpr int batched_lanes() {
  return KC == 64 ? lanes_per_query<64>() : FK_BATCHED_G;
}
__host__ __device__ constexpr int batched_class(int k) {
  return k <= 8 ? 0 : (k <= 16 ? 1 : (k <= 32 ? 2 : 3));
}
__host__ __device__ constexpr int batched_tile(int cls) {
  return kWave / (cls == 0   ? batched_lanes<8>()
                  : cls == 1 ? batched_lanes<16>()
                  : cls == 2 ? batched_lanes<32>()
                             : batched_lanes<64>());
}
static_assert(batched_class(64) == kBatchedClasses - 1, "batched classes");

// Candidate batch of the batched first pass (FK_BATCHED_CB_<KC>: the
// BatchSel of each class's kernel): the default of the K class, the large
// one for K class 32 (its LiDAR and room clouds ran 8 % faster with it
// alone, the 16-cloud batch 1-3 %; small batches for K classes 8/16 cost the
// batch 13 %, large ones changed nothing).
#ifndef FK_BATCHED_CB_8
#define FK_BATCHED_CB_8 0
#endif
#ifndef FK_BATCHED_CB_16
#define FK_BATCHED_CB_16 0
#endif
#ifndef FK_BATCHED_CB_32
#define FK_BATCHED_CB_32 2
#endif
template <int KC>
constexpr int batched_records() {
  return KC == 8    ? batch_records<8>(FK_BATCHED_CB_8)
         : KC == 16 ? batch_records<16>(FK_BATCHED_CB_16)
         : KC == 32 ? batch_records<32>(FK_BATCHED_CB_32)
                    : default_batch<KC>();
}

// Batched first pass of one K class: a persistent kernel whose waves pull
// (cloud, tile) work items of the clouds of that class from the class's own
// queue (tiles of kWave / G queries).  One kernel per class keeps each
// class's register budget and occupancy: a single kernel for every class
// ran all tiles at the budget of the largest (two waves per SIMD).
template <int KC>
__global__ void __launch_bounds__(64 * waves_per_block<KC>(), min_waves_per_eu<KC>())
k_batched(QueryArgs a, int cls) {
  constexpr int WPB = waves_per_block<KC>();
  constexpr int BS = buffer_slots<KC>();
  constexpr int G = batched_lanes<KC>();
  __shared__ uint64_t sbuf[WPB * BS * kWave];
  __shared__ uint64_t stab[WPB * stream_entries<KC>() * kWave];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
  uint64_t *lane_buf = sbuf + w * BS * kWave + lane;
  uint64_t *lane_tab = stab + w * stream_entries<KC>() * kWave + lane;
  const int32_t *prefix = a.item_cloud + cls * (a.num_clouds + 1);
  int32_t *head = a.scalars + kQueueHead + cls;
  const int nitems = a.scalars[kStatus] != 0 ? 0 : prefix[a.num_clouds];
  // the class's pending list, located once per wave (a read of the base
  // per tile, next to the queue heads' atomics, tripled the kernel's time)
  QueryArgs b = a;
  if (a.pend_class >= 0) {
    const int base = a.scalars[kPendBase + a.pend_class];
    b.pend_q = a.pend_q + base;
    b.pend_keys = a.pend_keys + (int64_t)base * a.k;
    b.nq = a.nq - base;
    b.pend_class = -1;
  }
  for (;;) {
    int item = 0;
    if (lane == 0) item = atomicAdd(head, 1);
    item = __shfl(item, 0, kWave);
    if (item >= nitems) break;
    const int c = item_to_cloud(prefix, a.num_clouds, item);
    const CloudInfo ci = a.clouds[c];
    const int qb = a.cell_start[ci.base];
    const int qe = a.cell_start[ci.base + table_entries(ci)];
    const int q0 = qb + (item - prefix[c]) * (kWave / G);
    run_tile<KC, G, false, batched_records<KC>()>(b, q0, qe, lane_buf, lane_tab);
  }
}

#if !FK_COLUMNS && FK_COMMON
// Per-class tile counts -> one exclusive prefix per K class over the clouds
// (one block): prefix[cls * (B + 1) + c] is cloud c's first item in class
// cls's queue (clouds of other classes contribute none).  Each class's
// pending list starts at scalars[kPendBase + cls], after the queries of
// the classes below it.  A cloud whose k lies outside [1, max_k] raises
// FK_ERR_K_RANGE, and then no tile runs.  qcnt: the counter set of class 0
// (the classes' sets follow it), cleared here.
__global__ void __launch_bounds__(1024)
k_items(const CloudInfo *clouds, const int32_t *cell_start, const int32_t *k_per_cloud,
        int32_t B, int32_t max_k, int32_t *prefix, int32_t *scalars, int32_t *qcnt) {
  __shared__ int wsum[kBatchedClasses][16];
  __shared__ int carry[kBatchedClasses];
  __shared__ int queries[kBatchedClasses];
  if (threadIdx.x < kBatchedClasses) carry[threadIdx.x] = queries[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // After a build that raised an error (an overflowing no-sync build still
  // writes every cloud's base, past the cell table) no cloud has a tile.
  const bool ok = scalars[kStatus] == 0;
  for (int c0 = 0; c0 < B; c0 += 1024) {
    const int c = c0 + threadIdx.x;
    int tiles = 0, cls = -1;
    if (ok && c < B) {
      const int k = k_per_cloud[c];
      if (k < 1 || k > max_k) {
        raise_status(scalars, FK_ERR_K_RANGE);
      } else {
        const CloudInfo ci = clouds[c];
        const int n = cell_start[ci.base + table_entries(ci)] - cell_start[ci.base];
        cls = batched_class(k);
        const int tile = batched_tile(cls);
        tiles = (n + tile - 1) / tile;
        if (n > 0) atomicAdd(&queries[cls], n);
      }
    }
    int v[kBatchedClasses];
#pragma unroll
    for (int j = 0; j < kBatchedClasses; ++j) {
      v[j] = cls == j ? tiles : 0;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v[j], o, kWave);
        if (lane >= o) v[j] += u;
      }
      if (lane == 63) wsum[j][w] = v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kBatchedClasses; ++j) {
      int before = carry[j];
      for (int i = 0; i < w; ++i) before += wsum[j][i];
      if (c < B) prefix[j * (B + 1) + c] = before + v[j] - (cls == j ? tiles : 0);
      v[j] += before;
    }
    __syncthreads();
    if (threadIdx.x == 1023) {
#pragma unroll
      for (int j = 0; j < kBatchedClasses; ++j) carry[j] = v[j];
    }
    __syncthreads();
  }
  if (threadIdx.x < kBatchedClasses) {
    prefix[threadIdx.x * (B + 1) + B] = carry[threadIdx.x];
    scalars[kQueueHead + threadIdx.x] = 0;
    int below = 0;
    for (int j = 0; j < (int)threadIdx.x; ++j) below += queries[j];
    scalars[kPendBase + threadIdx.x] = below;
  }
  if (threadIdx.x < kCntStride * kBatchedClasses) qcnt[threadIdx.x] = 0;
}
#endif

// Sum over queries of the candidate count |C(q)| of the block of
// half-width R (the reference's traverse statistic,
// FlashKNN/functions/FlashKnnWrapper.py:351-352).
// One atomic per block (kCountBlock threads) instead of one per wave on
// the same counter (per-cloud sums on the 98k-point room: 22 -> 8 us).
constexpr int kCountBlock = 1024;

__device__ __forceinline__ unsigned long long query_candidates(const QueryArgs &a,
                                                               int qi, int &cloud) {
  const int4 qc = a.qcell[qi];
  cloud = qc.w;
  const CloudInfo ci = a.clouds[qc.w];
  unsigned long long c = 0;
  for (int dx = -a.R; dx <= a.R; ++dx)
    for (int dy = -a.R; dy <= a.R; ++dy) {
      int b, e;
      column_range(ci, a.cell_start, a.scell, qc.x + dx, qc.y + dy, qc.z - a.R,
                   qc.z + a.R, b, e);
      c += (unsigned long long)(e - b);
    }
  return c;
}

#if FK_COMMON
__global__ void __launch_bounds__(kCountBlock)
k_candidate_count(const QueryArgs a, unsigned long long *total) {
  __shared__ unsigned long long wsum[kCountBlock / 64];
  const int qi = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long c = 0;
  int cloud;
  if (qi < a.nq && a.scalars[kStatus] == 0) c = query_candidates(a, qi, cloud);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int j = 0; j < (int)(blockDim.x >> 6); ++j) t += wsum[j];
    if (t) atomicAdd(total, t);
  }
}
#endif  // FK_COMMON

// Largest block candidate count of every super-tile of kPlanTile = kWave
// consecutive queries (one wave per super-tile).
#if FK_COMMON
__global__ void __launch_bounds__(256) k_tile_cost(const QueryArgs a, int32_t *tile_cost) {
  const int qi = blockIdx.x * blockDim.x + threadIdx.x;
  int c = 0;
  if (qi < a.nq && a.scalars[kStatus] == 0) {
    int cloud;
    const unsigned long long v = query_candidates(a, qi, cloud);
    c = v < (unsigned long long)INT_MAX ? (int)v : INT_MAX;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c = max(c, __shfl_xor(c, o, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0 && qi < a.nq) tile_cost[qi / kPlanTile] = c;
}
#endif  // FK_COMMON

constexpr int kPlanMaxTiles = 8192;  // super-tiles cached in LDS by k_plan

constexpr int kPlanSearch = 7;       // bisection steps over the 2^7 budgets

// One lane per query (64 queries per wave) measured about 1.4x slower per
// candidate than two lanes per query (a lane's buffer fills from one query's
// stream, with no threshold shared in the group): a super-tile takes one
// lane only if its cost with that factor fits the budget.
#ifndef FK_PLAN_ONE_LANE_COST
#define FK_PLAN_ONE_LANE_COST 2.0f
#endif
// K class 32 (its lists fill a lane's registers either way, and two lanes
// per query double the merge work) takes one lane at cost factor 1: room
// 98k first pass 85.3 -> 81.7 us, LiDAR sweep 92k 110.8 -> 106.7 us; for K
// classes 8/16 factors of 1 and 1.5 were slower (room K=16 43.5 -> 48.5 /
// 44.3 us, LiDAR 46.4 -> 55.5 / 50.8 us).
#ifndef FK_PLAN_ONE_LANE_COST_32
#define FK_PLAN_ONE_LANE_COST_32 1.0f
#endif
constexpr int kPlanGmin = 1;

// Lanes per query of a super-tile whose largest count is `cost` under the
// per-lane budget L: 1 if cost * FK_PLAN_ONE_LANE_COST <= L, else the
// smallest power of two in [2, 8] with cost <= G * L.
__device__ __forceinline__ int plan_lanes(int cost, float L, float one_lane_cost) {
  if ((float)cost * one_lane_cost <= L) return 1;
  int g = 2;
  while (g < 8 && (float)cost > (float)g * L) g <<= 1;
  return g;
}

// Budget level j (0..127): L = 4 * 2^(j / 8) candidates per lane.
__device__ __forceinline__ float plan_budget(int j) { return 4.f * exp2f(0.125f * (float)j); }

// One block: the smallest budget whose waves fit `capacity`, then the
// per-wave plan, super-tiles with more lanes first (they are the longest
// waves: started first, they finish with the rest).  Unused plan entries
// get (nq, 1) and exit at once.  The waves go to the XCDs one by one
// (blocks are dealt round-robin over them): slabs of consecutive
// super-tiles per XCD cut the L2 fetch by up to half but made the pass no
// faster on the room and slower on a LiDAR sweep and at K = 32.
#if FK_COMMON
__global__ void __launch_bounds__(1024)
k_plan(const int32_t *tile_cost, int32_t ntiles, int32_t nq, int32_t capacity, float olc,
       const int32_t *scalars, int2 *plan) {
  __shared__ int cost[kPlanMaxTiles];
  __shared__ int red[16];
  __shared__ int cls_n[4], carry[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const bool ok = scalars[kStatus] == 0;
  if (!ok) {
    for (int i = t; i < capacity; i += 1024) plan[i] = make_int2(nq, 1);
    return;
  }
  for (int i = t; i < ntiles; i += 1024) cost[i] = tile_cost[i];
  __syncthreads();
  auto block_sum = [&](int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    if (lane == 0) red[w] = v;
    __syncthreads();
    int s = 0;
    for (int j = 0; j < 16; ++j) s += red[j];
    __syncthreads();
    return s;
  };
  // waves(L) does not grow with L: bisect the budget levels
  int lo = 0, hi = (1 << kPlanSearch) - 1;  // hi always fits (8192 per lane)
  for (int it = 0; it < kPlanSearch; ++it) {
    const int mid = (lo + hi) >> 1;
    const float L = plan_budget(mid);
    int v = 0;
    for (int i = t; i < ntiles; i += 1024) v += plan_lanes(cost[i], L, olc);
    if (block_sum(v) <= capacity) hi = mid; else lo = mid + 1;
  }
  // the last level puts every super-tile on kPlanGmin lanes per query
  // (waves = kPlanGmin * super-tiles <= capacity, which the host checks)
  const float L = hi == (1 << kPlanSearch) - 1 ? INFINITY : plan_budget(hi);
  // super-tiles per class (G = 8, 4, 2, 1): the four counts in 16-bit
  // fields of two sums (ntiles <= kPlanMaxTiles)
  static_assert(kPlanMaxTiles < (1 << 16), "16-bit class counts");
  {
    int v01 = 0, v23 = 0;
    for (int i = t; i < ntiles; i += 1024) {
      const int g = plan_lanes(cost[i], L, olc);
      v01 += g == 8 ? 1 : (g == 4 ? 1 << 16 : 0);
      v23 += g == 2 ? 1 : (g == 1 ? 1 << 16 : 0);
    }
    const int s01 = block_sum(v01), s23 = block_sum(v23);
    if (t == 0) {
      cls_n[0] = s01 & 0xffff;
      cls_n[1] = s01 >> 16;
      cls_n[2] = s23 & 0xffff;
      cls_n[3] = s23 >> 16;
      int base = 0;
      for (int c = 0; c < 4; ++c) {
        carry[c] = base;  // running first wave of the next super-tile of class c
        base += cls_n[c] * (8 >> c);
      }
    }
  }
  __syncthreads();
  __shared__ int wcnt[4][16];
  for (int i0 = 0; i0 < ntiles; i0 += 1024) {
    const int i = i0 + t;
    const int g = i < ntiles ? plan_lanes(cost[i], L, olc) : 0;
    const int c = g == 8 ? 0 : (g == 4 ? 1 : (g == 2 ? 2 : 3));
    int pos = 0;
    for (int cc = 0; cc < 4; ++cc) {
      const uint64_t m = __ballot(g != 0 && c == cc);
      if (lane == 0) wcnt[cc][w] = __popcll(m);
      if (g != 0 && c == cc) pos = __popcll(m & ((1ull << lane) - 1));
    }
    __syncthreads();
    if (g != 0) {
      int before = 0;  // super-tiles of class c ahead of this one in the chunk
      for (int j = 0; j < w; ++j) before += wcnt[c][j];
      const int first = carry[c] + (before + pos) * g;
      for (int p = 0; p < g; ++p)
        if (first + p < capacity)  // always (the waves fit by construction)
          plan[first + p] = make_int2(i * kPlanTile + p * (kPlanTile / g), g);
    }
    __syncthreads();
    if (t < 4) {
      int tot = 0;
      for (int j = 0; j < 16; ++j) tot += wcnt[t][j];
      carry[t] += tot * (8 >> t);
    }
    __syncthreads();
  }
  int used = 0;
  for (int c = 0; c < 4; ++c) used += cls_n[c] * (8 >> c);
  for (int i = used + t; i < capacity; i += 1024) plan[i] = make_int2(nq, 1);
}
#endif  // FK_COMMON

// Candidate count |C(q)| of every query, written to its output row.
#if FK_COMMON
__global__ void k_point_candidates(const QueryArgs a, int32_t *out) {
  const int qi = blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= a.nq || a.scalars[kStatus] != 0) return;
  const int4 qc = a.qcell[qi];
  const CloudInfo ci = a.clouds[qc.w];
  long long c = 0;
  for (int dx = -a.R; dx <= a.R; ++dx)
    for (int dy = -a.R; dy <= a.R; ++dy) {
      int b, e;
      column_range(ci, a.cell_start, a.scell, qc.x + dx, qc.y + dy, qc.z - a.R,
                   qc.z + a.R, b, e);
      c += e - b;
    }
  out[__float_as_int(a.qrec[qi].w)] = (int32_t)(c < INT_MAX ? c : INT_MAX);
}
#endif  // FK_COMMON

// Per-cloud candidate sums (self queries, block half-width a.R).  A block
// whose queries all lie in one cloud (queries are sorted by cloud, so all
// but the blocks at cloud boundaries) adds once; the others per wave, or per
// lane where a wave spans clouds.
#if FK_COMMON
__global__ void __launch_bounds__(kCountBlock)
k_cand_per_cloud(const QueryArgs a, unsigned long long *acc) {
  __shared__ unsigned long long wsum[kCountBlock / 64];
  __shared__ int wcloud[kCountBlock / 64];
  const int qi = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = qi < a.nq && a.scalars[kStatus] == 0;
  unsigned long long c = 0;
  int cloud = -1;
  if (ok) c = query_candidates(a, qi, cloud);
  const int w = threadIdx.x >> 6;
  const uint64_t m = __ballot(ok);
  // wave cloud: -1 empty, -2 mixed (its lanes already added), else the cloud
  int wc = -1;
  if (m != 0ull) {
    const int c0 = __shfl(cloud, (int)__builtin_ctzll(m), kWave);
    if (__ballot(ok && cloud != c0) == 0ull) {
      wc = c0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, kWave);
    } else {
      wc = -2;
      if (ok && c) atomicAdd(acc + cloud, c);
    }
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    wsum[w] = c;
    wcloud[w] = wc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    int cur = -1;
    unsigned long long run = 0;
    for (int j = 0; j < nw; ++j) {
      const int cj = wcloud[j];
      if (cj < 0) continue;
      if (cj != cur) {
        if (cur >= 0 && run) atomicAdd(acc + cur, run);
        cur = cj;
        run = 0;
      }
      run += wsum[j];
    }
    if (cur >= 0 && run) atomicAdd(acc + cur, run);
  }
}
#endif  // FK_COMMON

// side *= clamp((target_ratio * k / mean candidates)^(1/exponent)), bounded
// so that the rebuilt cell box stays below ~64 cells per point.
#if FK_COMMON
__global__ void k_adapt(const CloudInfo *clouds, int32_t B,
                        const unsigned long long *acc, const int32_t *kpc,
                        int32_t k, float target_ratio, const float *ratio_pc,
                        float exponent, float *side, float *inv) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= B) return;
  const CloudInfo ci = clouds[c];
  const double n = (double)ci.num_points;
  if (n <= 0.0 || acc[c] == 0ull) return;
  const double mean = (double)acc[c] / n;
  const double ratio = ratio_pc ? (double)ratio_pc[c] : (double)target_ratio;
  const double target = ratio * (double)(kpc ? kpc[c] : k);
  double f = pow(target / mean, 1.0 / (double)exponent);
  f = fmin(fmax(f, 0.25), 4.0);
  const double cells = (double)ci.nx * (double)ci.ny * (double)ci.nz;
  f = fmax(f, cbrt(cells / (64.0 * n)));
  const float s = (float)((double)side[c] * f);
  side[c] = s;
  inv[c] = (float)(1.0 / (double)s);
}
#endif  // FK_COMMON

#if FK_COMMON
__global__ void k_zero_u64(unsigned long long *p, int32_t n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}
#endif  // FK_COMMON

#if FK_COMMON
__global__ void k_reset_query_counters(int32_t *qcnt) {
  if (threadIdx.x < kCntStride) qcnt[threadIdx.x] = 0;
}
#endif  // FK_COMMON

// Fixed grid (graph-capturable): pending queries are read on the device.
// Up to kExpandWaves waves, static striding over the pending list: two
// waves per SIMD, each finishing a few queries one after another, beat
// more waves that share their SIMDs' issue (LiDAR sweep, 7,614 pending at
// K=16, whole query: 1,024 / 2,048 / 3,072 / 4,096 / 8,192 waves 124 /
// 108 / 112 / 121 / 159 us; the 16-cloud batched replay 0.52 / 0.468 /
// 0.467 / 0.476 ms; one wave per pending query, 16,384 waves, slower
// still).
#ifndef FK_EXPAND_WAVES
#define FK_EXPAND_WAVES 2048
#endif
constexpr int kExpandWaves = FK_EXPAND_WAVES;

[[maybe_unused]] unsigned expand_blocks(int nq) {
  const int waves = nq < kExpandWaves ? nq : kExpandWaves;
  const int b = (waves + kCoopWaves - 1) / kCoopWaves;
  return (unsigned)(b < 1 ? 1 : b);
}

template <int KC, int G, int CB = default_batch<KC>()>
void launch_first(const QueryArgs &a, hipStream_t s) {
  constexpr int WPB = waves_per_block<KC>();
  constexpr int QPW = kWave / G;  // queries per wave
  const int64_t waves = ((int64_t)a.nq + QPW - 1) / QPW;
  const unsigned nb = (unsigned)((waves + WPB - 1) / WPB);
  hipLaunchKernelGGL((k_query<KC, G, CB>), dim3(nb), dim3(64 * WPB), 0, s, a);
}

// f(std::integral_constant<int, CB>) for the candidate batch that sel
// selects.  The alternatives exist for the K classes 8-32 of the dense
// layout; the column layout and K class 64 keep the default.
template <int KC, class F>
void with_batch(int sel, F &&f) {
#if !FK_COLUMNS
  if constexpr (KC <= 32) {
    if (sel == kBatchSmall) {
      f(std::integral_constant<int, batch_records<KC>(kBatchSmall)>());
      return;
    }
    if (sel == kBatchLarge) {
      f(std::integral_constant<int, batch_records<KC>(kBatchLarge)>());
      return;
    }
  }
#endif
  f(std::integral_constant<int, default_batch<KC>()>());
}

// A launch of few queries is bound by the latency of its longest scans, not
// by throughput: below kFillWaves waves (8 per CU) the group per query is
// widened (up to 16 lanes, each scanning every 16th candidate) until the
// launch has that many waves.  The merged lists, and so the results, do not
// depend on the group size.
#ifndef FK_FILL_WAVES
#define FK_FILL_WAVES 2048
#endif
constexpr int64_t kFillWaves = FK_FILL_WAVES;

template <int KC, int G, int CB>
void launch_grouped(const QueryArgs &a, hipStream_t s) {
  const int64_t lanes = (int64_t)a.nq * G;
  if constexpr (G < 16 && buffer_slots<KC>() >= KC) {
    if (lanes < kFillWaves * kWave) {
      launch_grouped<KC, 2 * G, CB>(a, s);
      return;
    }
  }
  launch_first<KC, G, CB>(a, s);
}

// First pass of K class KC: the balanced plan when the launch has one,
// else uniform tiles (one lane per query with FK_QUERY_ONE_LANE, widened
// groups for small launches).
template <int KC>
void launch_plain(const QueryArgs &a, hipStream_t s) {
  constexpr int G = lanes_per_query<KC>();
  if constexpr (KC <= 32) {
    if (a.plan) {
      constexpr int WPB = waves_per_block<KC>();
      with_batch<KC>(a.batch_sel, [&](auto cb) {
        hipLaunchKernelGGL((k_query_planned<KC, decltype(cb)::value>),
                           dim3(a.plan_waves / WPB), dim3(64 * WPB), 0, s, a);
      });
      return;
    }
  }
  if (G > 1 && a.one_lane) {
    with_batch<KC>(a.batch_sel,
                   [&](auto cb) { launch_first<KC, 1, decltype(cb)::value>(a, s); });
  } else if constexpr (G > 1 || KC <= 32) {
    // the candidate batch of the hint; small launches widen their groups
    with_batch<KC>(a.batch_sel,
                   [&](auto cb) { launch_grouped<KC, G, decltype(cb)::value>(a, s); });
  } else {
    launch_first<KC, G>(a, s);
  }
}

}  // namespace

// tags of the cached occupancy queries
template <int KC>
struct BatchedTag {};
template <int KC, int CB>
struct PlannedTag {};

// ------------------------------------------------- per K class launchers
// (defined and explicitly instantiated only in the unit of their class: the
// other units see the declarations and do not compile the class's kernels)
#if FK_QPART == 0 || FK_QPART >= 8

template <int KC>
hipError_t FK_LAYOUT(launch_first_pass)(const QueryArgs &a, hipStream_t s) {
  launch_plain<KC>(a, s);
  return hipGetLastError();
}

// Persistent: as many resident workgroups as fit pull the class's tiles.
template <int KC>
hipError_t FK_LAYOUT(launch_batched_class)(const QueryArgs &a, int cls, hipStream_t s) {
  constexpr int WPB = waves_per_block<KC>();
  int cus = 0, nb = 0;
  hipError_t e = device_cus(&cus);
  if (e == hipSuccess)
    e = cached_per_device<BatchedTag<KC>>(
        [](int, int *r) {
          return hipOccupancyMaxActiveBlocksPerMultiprocessor(r, k_batched<KC>, 64 * WPB, 0);
        },
        &nb);
  if (e != hipSuccess) return e;
  const unsigned grid = (unsigned)((nb < 1 ? 1 : nb) * cus);
  hipLaunchKernelGGL(k_batched<KC>, dim3(grid), dim3(64 * WPB), 0, s, a, cls);
  return hipGetLastError();
}

#if !FK_COLUMNS
template <int KC, int CB>
hipError_t planned_blocks_cb(int *nb) {
  constexpr int kThreads = 64 * waves_per_block<KC>();
  return cached_per_device<PlannedTag<KC, CB>>(
      [](int, int *r) {
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(r, k_query_planned<KC, CB>,
                                                            kThreads, 0);
      },
      nb);
}

template <int KC>
hipError_t planned_blocks(int sel, int *nb) {
  hipError_t e = hipSuccess;
  with_batch<KC>(sel, [&](auto cb) { e = planned_blocks_cb<KC, decltype(cb)::value>(nb); });
  return e;
}
#define FK_PLANNED_BLOCKS(KC) template hipError_t planned_blocks<KC>(int, int *);
#else
#define FK_PLANNED_BLOCKS(KC)
#endif

#if !FK_COLUMNS
template <int KC>
hipError_t launch_adaptive_query(const AdaptiveArgs &a, hipStream_t s) {
  constexpr int WPB = waves_per_block<KC>();
  constexpr int QPW = kWave / kAdG;
  if (a.n <= 0) return hipSuccess;
  const int64_t waves = ((int64_t)a.n + QPW - 1) / QPW;
  const unsigned nb = (unsigned)((waves + WPB - 1) / WPB);
  hipLaunchKernelGGL(k_adaptive_query<KC>, dim3(nb), dim3(64 * WPB), 0, s, a);
  return hipGetLastError();
}
#define FK_ADAPTIVE_UNIT(KC) \
  template hipError_t launch_adaptive_query<KC>(const AdaptiveArgs &, hipStream_t);
#else
#define FK_ADAPTIVE_UNIT(KC)
#endif

#define FK_CLASS_UNIT(KC)                                                        \
  template hipError_t FK_LAYOUT(launch_first_pass)<KC>(const QueryArgs &,        \
                                                       hipStream_t);             \
  template hipError_t FK_LAYOUT(launch_batched_class)<KC>(const QueryArgs &, int, \
                                                          hipStream_t);          \
  FK_PLANNED_BLOCKS(KC)                                                          \
  FK_ADAPTIVE_UNIT(KC)
#if FK_CLASS(8)
FK_CLASS_UNIT(8)
#endif
#if FK_CLASS(16)
FK_CLASS_UNIT(16)
#endif
#if FK_CLASS(32)
FK_CLASS_UNIT(32)
#endif
#if FK_CLASS(64)
template hipError_t FK_LAYOUT(launch_first_pass)<64>(const QueryArgs &, hipStream_t);
template hipError_t FK_LAYOUT(launch_batched_class)<64>(const QueryArgs &, int, hipStream_t);
FK_ADAPTIVE_UNIT(64)
#endif
#endif  // class units

// ------------------------------------------------------- dispatching launchers
#if FK_COMMON

#if !FK_COLUMNS
int kclass_of(int k) {
  return k <= 8 ? 8 : (k <= 16 ? 16 : (k <= 32 ? 32 : (k <= 64 ? 64 : 128)));
}
#define FK_FORWARD(call) \
  if (a.columns) return call
#else
#define FK_FORWARD(call)
#endif

hipError_t FK_LAYOUT(launch_query)(const QueryArgs &a, hipStream_t s) {
  FK_FORWARD(launch_query_columns(a, s));
  if (a.passes & 1)
    hipLaunchKernelGGL(k_reset_query_counters, dim3(1), dim3(64), 0, s, a.qcnt);
  if (a.nq <= 0) return hipGetLastError();
  if (a.k > kMaxLaneK) {
    if (a.passes & 2) return FK_LAYOUT(launch_wide)(a, s);
    return hipGetLastError();
  }
  if (a.passes & 2) {
    hipError_t e;
    switch (kclass_of(a.k)) {
      case 8: e = FK_LAYOUT(launch_first_pass)<8>(a, s); break;
      case 16: e = FK_LAYOUT(launch_first_pass)<16>(a, s); break;
      case 32: e = FK_LAYOUT(launch_first_pass)<32>(a, s); break;
      default: e = FK_LAYOUT(launch_first_pass)<64>(a, s); break;
    }
    if (e != hipSuccess) return e;
  }
  if (a.exact && (a.passes & 4)) return FK_LAYOUT(launch_expand)(a, s);
  return hipGetLastError();
}

hipError_t FK_LAYOUT(launch_plan)(const QueryArgs &a, int32_t capacity,
                                  int32_t *tile_cost, int2 *plan, hipStream_t s) {
  FK_FORWARD(launch_plan_columns(a, capacity, tile_cost, plan, s));
  const int64_t ntiles = ((int64_t)a.nq + kPlanTile - 1) / kPlanTile;
  if (a.nq <= 0 || ntiles * kPlanGmin > capacity || ntiles > kPlanMaxTiles)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tile_cost, dim3((a.nq + 255) / 256), dim3(256), 0, s, a, tile_cost);
  hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, s, tile_cost, (int32_t)ntiles, a.nq,
                     capacity, a.k > 16 ? FK_PLAN_ONE_LANE_COST_32 : FK_PLAN_ONE_LANE_COST,
                     a.scalars, plan);
  return hipGetLastError();
}

#if !FK_COLUMNS
// Waves a plan may use beyond one resident round, percent (measured: any
// over-subscription was slower than fitting the round).
#ifndef FK_PLAN_OVER
#define FK_PLAN_OVER 0
#endif
constexpr int kPlanOver = FK_PLAN_OVER;

hipError_t planned_capacity(int kc, int batch_sel, int32_t *waves) {
  int cus = 0, nb = 0;
  hipError_t e = device_cus(&cus);
  if (e != hipSuccess) return e;
  int wpb = 0;
  switch (kc) {
    case 8:
      e = planned_blocks<8>(batch_sel, &nb);
      wpb = 4;
      break;
    case 16:
      e = planned_blocks<16>(batch_sel, &nb);
      wpb = 4;
      break;
    case 32:
      e = planned_blocks<32>(batch_sel, &nb);
      wpb = 2;
      break;
    default:
      return hipErrorInvalidValue;
  }
  if (e != hipSuccess) return e;
  static_assert(waves_per_block<8>() == 4 && waves_per_block<16>() == 4 &&
                    waves_per_block<32>() == 2,
                "planned capacity: waves per block");
  // the plan may exceed one resident round by kPlanOver percent: the
  // lightest super-tiles, planned last, start when the first waves retire
  const int64_t resident = (int64_t)(nb < 1 ? 1 : nb) * cus * wpb;
  *waves = (int32_t)((resident * (100 + kPlanOver) / 100) / wpb * wpb);
  return hipSuccess;
}
#endif

hipError_t FK_LAYOUT(launch_candidate_count)(const QueryArgs &a,
                                             unsigned long long *total,
                                             hipStream_t s) {
  FK_FORWARD(launch_candidate_count_columns(a, total, s));
  if (a.nq <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_candidate_count, dim3((a.nq + kCountBlock - 1) / kCountBlock),
                     dim3(kCountBlock), 0, s, a, total);
  return hipGetLastError();
}

hipError_t FK_LAYOUT(launch_point_candidates)(const QueryArgs &a, int32_t *out,
                                              hipStream_t s) {
  FK_FORWARD(launch_point_candidates_columns(a, out, s));
  if (a.nq <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_point_candidates, dim3((a.nq + 255) / 256), dim3(256), 0,
                     s, a, out);
  return hipGetLastError();
}

hipError_t FK_LAYOUT(launch_cloud_candidates)(const QueryArgs &a,
                                              unsigned long long *acc,
                                              hipStream_t s) {
  FK_FORWARD(launch_cloud_candidates_columns(a, acc, s));
  const int B = a.num_clouds;
  hipLaunchKernelGGL(k_zero_u64, dim3((B + 255) / 256), dim3(256), 0, s, acc, B);
  if (a.nq > 0)
    hipLaunchKernelGGL(k_cand_per_cloud, dim3((a.nq + kCountBlock - 1) / kCountBlock),
                       dim3(kCountBlock), 0, s, a, acc);
  return hipGetLastError();
}

hipError_t FK_LAYOUT(launch_adapt_cell)(const QueryArgs &a, unsigned long long *acc,
                                        const int32_t *k_per_cloud, int32_t k,
                                        float target_ratio, const float *ratio_pc,
                                        float exponent, float *side_io,
                                        float *inv_io, hipStream_t s) {
  FK_FORWARD(launch_adapt_cell_columns(a, acc, k_per_cloud, k, target_ratio,
                                       ratio_pc, exponent, side_io, inv_io, s));
  const int B = a.num_clouds;
  hipLaunchKernelGGL(k_zero_u64, dim3((B + 255) / 256), dim3(256), 0, s, acc, B);
  if (a.nq > 0)
    hipLaunchKernelGGL(k_cand_per_cloud, dim3((a.nq + kCountBlock - 1) / kCountBlock),
                       dim3(kCountBlock), 0, s, a, acc);
  hipLaunchKernelGGL(k_adapt, dim3((B + 255) / 256), dim3(256), 0, s, a.clouds, B,
                     acc, k_per_cloud, k, target_ratio, ratio_pc, exponent, side_io,
                     inv_io);
  return hipGetLastError();
}

#if !FK_COLUMNS
hipError_t launch_items(const CloudInfo *clouds, const int32_t *cell_start,
                        const int32_t *k_per_cloud, int32_t B, int32_t max_k,
                        int32_t *prefix, int32_t *scalars, int32_t *qcnt,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_items, dim3(1), dim3(1024), 0, s, clouds, cell_start, k_per_cloud,
                     B, max_k, prefix, scalars, qcnt);
  return hipGetLastError();
}
#endif  // !FK_COLUMNS

hipError_t FK_LAYOUT(launch_batched_first)(const QueryArgs &a, int cls, hipStream_t s) {
  FK_FORWARD(launch_batched_first_columns(a, cls, s));
  switch (cls) {
    case 0: return FK_LAYOUT(launch_batched_class)<8>(a, cls, s);
    case 1: return FK_LAYOUT(launch_batched_class)<16>(a, cls, s);
    case 2: return FK_LAYOUT(launch_batched_class)<32>(a, cls, s);
    case 3: return FK_LAYOUT(launch_batched_class)<64>(a, cls, s);
    default: return hipErrorInvalidValue;
  }
}

#endif  // FK_COMMON

#if FK_EXPAND_PART
hipError_t FK_LAYOUT(launch_expand)(const QueryArgs &a, hipStream_t s) {
  if (a.exact && a.nq > 0)
    hipLaunchKernelGGL(k_expand_coop, dim3(expand_blocks(a.nq)), dim3(64 * kCoopWaves), 0,
                       s, a);
  return hipGetLastError();
}
#endif

#if FK_WIDE_PART
hipError_t FK_LAYOUT(launch_wide)(const QueryArgs &a, hipStream_t s) {
  const int64_t b = ((int64_t)a.nq + kCoopWaves - 1) / kCoopWaves;
  const unsigned nb = (unsigned)(b < 16384 ? b : 16384);
  hipLaunchKernelGGL(k_query_wide, dim3(nb), dim3(64 * kCoopWaves), 0, s, a);
  return hipGetLastError();
}
#endif

}  // namespace fk
