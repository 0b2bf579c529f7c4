// Backward of the gathered squared distances and IDW interpolation (gfx950).
//
// No counterpart in the reference: FlashKNN.query is @torch.no_grad
// (FlashKNN/functions/FlashKnnWrapper.py:165) and pointops' interpolation
// backward scatters features with float atomicAdd
// (Pointcept/libs/pointops/src/interpolation/interpolation_cuda_kernel.cu:31),
// whose summation order, and so whose rounding, changes from run to run.
// Here every support-side gradient is a segmented sum over a reverse CSR of
// the neighbour table (entries e = i*k + j grouped by idx[e], each segment
// sorted by e), so the sum order is fixed and results are bitwise
// reproducible.
#include "fk_internal.h"

namespace fk {

hipError_t launch_scan_host(const int32_t *in, int32_t *out, int64_t n,
                            int32_t *part, hipStream_t s);
int64_t scan_parts_needed(int64_t n);

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int lane_value(int v, int j) {
  return __builtin_amdgcn_readlane(v, j);
}
__device__ __forceinline__ float lane_value(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

inline unsigned nblocks(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  return (unsigned)(b < 1 ? 1 : b);
}

struct Workspace {
  int32_t *counts;  // [ns]
  int32_t *start;   // [ns + 1]
  int32_t *part;    // scan partials
  int32_t *rank;    // [E] rank in the segment, then the sorted segments
  int32_t *slots;   // [E] placed entries (segment order unsorted)
};

inline int64_t align_up(int64_t v) { return (v + 255) & ~int64_t(255); }

Workspace carve(void *ws, int64_t ns, int64_t E) {
  char *p = (char *)ws;
  Workspace w;
  w.counts = (int32_t *)p; p += align_up(4 * ns);
  w.start = (int32_t *)p; p += align_up(4 * (ns + 1));
  w.part = (int32_t *)p; p += align_up(4 * scan_parts_needed(ns));
  w.rank = (int32_t *)p; p += align_up(4 * E);
  w.slots = (int32_t *)p;
  return w;
}

__global__ void k_zero(int32_t *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void k_rc_count(const int64_t *idx, int64_t E, int64_t ns,
                           int32_t *counts, int32_t *rank) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int64_t m = idx[e];
  rank[e] = (m >= 0 && m < ns) ? atomicAdd(counts + m, 1) : -1;
}

__global__ void k_rc_place(const int64_t *idx, int64_t E, const int32_t *start,
                           const int32_t *rank, int32_t *slots) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int32_t r = rank[e];
  if (r >= 0) slots[start[idx[e]] + r] = (int32_t)e;
}

// Sort every segment by entry id (fixes the summation order): one wave per
// segment; each lane takes one entry of a 64-entry chunk and counts the
// segment's entries below it (ids are distinct, so the counts are the
// ranks), reading the segment from LDS when it fits.
constexpr int kRankWaves = 4;
constexpr int kRankStage = 512;  // entries staged in LDS per wave

__global__ void __launch_bounds__(64 * kRankWaves)
k_rc_rank(const int32_t *start, int64_t ns, const int32_t *unsorted,
          int32_t *sorted) {
  __shared__ int32_t stage[kRankWaves * kRankStage];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t m = (int64_t)blockIdx.x * kRankWaves + w;
  if (m >= ns) return;  // wave-uniform
  const int32_t b = start[m], L = start[m + 1] - b;
  const int32_t *seg = unsorted + b;
  int32_t *st = stage + w * kRankStage;
  const bool staged = L <= kRankStage;
  if (staged)
    for (int t = lane; t < L; t += 64) st[t] = seg[t];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int32_t *src = staged ? st : seg;
  for (int c0 = 0; c0 < L; c0 += 64) {
    const bool have = c0 + lane < L;
    const int32_t x = have ? src[c0 + lane] : 0;
    int r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    int t = 0;
    for (; t + 4 <= L; t += 4) {
      r0 += src[t] < x;
      r1 += src[t + 1] < x;
      r2 += src[t + 2] < x;
      r3 += src[t + 3] < x;
    }
    for (; t < L; ++t) r0 += src[t] < x;
    if (have) sorted[b + r0 + r1 + r2 + r3] = x;
  }
}

__global__ void k_dist_bw_q(const float *q, int64_t nq, const float *p,
                            int64_t ns, const int64_t *idx, int32_t k,
                            const float *g, float *gq) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const float qx = q[3 * i], qy = q[3 * i + 1], qz = q[3 * i + 2];
  float ax = 0.f, ay = 0.f, az = 0.f;
  for (int j = 0; j < k; ++j) {
    const int64_t m = idx[i * k + j];
    if (m < 0 || m >= ns) continue;
    const float c = 2.f * g[i * k + j];
    ax += c * (qx - p[3 * m]);
    ay += c * (qy - p[3 * m + 1]);
    az += c * (qz - p[3 * m + 2]);
  }
  gq[3 * i] = ax;
  gq[3 * i + 1] = ay;
  gq[3 * i + 2] = az;
}

// One wave per support m: each 64-entry chunk of its (ascending) segment is
// loaded one entry per lane, every lane forms its entry's term, and the terms
// are added in entry order by reading them back lane by lane, so the sum is
// the same as a sequential loop's (a thread per support walked ~50-70
// dependent entry loads: ~45 us whatever the size).
__global__ void __launch_bounds__(256)
k_dist_bw_p(const float *q, const float *p, int64_t ns, int32_t k,
            const float *g, const int32_t *start, const int32_t *slots,
            float *gp) {
  const int64_t m = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (m >= ns) return;  // wave-uniform
  const float px = p[3 * m], py = p[3 * m + 1], pz = p[3 * m + 2];
  float ax = 0.f, ay = 0.f, az = 0.f;
  const int32_t b = start[m], end = start[m + 1];
  for (int32_t c0 = b; c0 < end; c0 += 64) {
    const int n = end - c0 < 64 ? end - c0 : 64;
    float tx = 0.f, ty = 0.f, tz = 0.f;
    if (lane < n) {
      const int32_t e = slots[c0 + lane];
      const int64_t i = e / k;
      const float c = 2.f * g[e];
      tx = c * (q[3 * i] - px);
      ty = c * (q[3 * i + 1] - py);
      tz = c * (q[3 * i + 2] - pz);
    }
    for (int u = 0; u < n; ++u) {
      ax -= lane_value(tx, u);
      ay -= lane_value(ty, u);
      az -= lane_value(tz, u);
    }
  }
  if (lane == 0) {
    gp[3 * m] = ax;
    gp[3 * m + 1] = ay;
    gp[3 * m + 2] = az;
  }
}

// out[i, c] = sum_j w[i,j] feat[idx[i,j], c]  (thread per (i, c)).
__global__ void k_idw_fwd(const float *feat, int64_t ns, int32_t C,
                          const int64_t *idx, const float *w, int64_t nq,
                          int32_t k, float *out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nq * C) return;
  const int64_t i = t / C;
  const int c = (int)(t - i * C);
  float acc = 0.f;
  for (int j = 0; j < k; ++j) {
    const int64_t m = idx[i * k + j];
    if (m < 0 || m >= ns) continue;
    acc += w[i * k + j] * feat[m * C + c];
  }
  out[t] = acc;
}

// grad_feat[m, c] = sum over segment m (entry order) of w[e] grad_out[e/k, c].
// grad_feat[m, c] = sum over the entries e of support m (ascending) of
// w[e] * grad_out[e / k, c].  One wave per (support, 64-channel chunk), lane
// = channel: the segment's entry ids and weights are loaded 64 at a time
// (one per lane) and broadcast, so the loop only gathers grad_out rows
// (coalesced over the channels), eight in flight; the sum runs in entry
// order.
__global__ void __launch_bounds__(256)
k_idw_bw_feat(int64_t ns, int32_t C, const float *w, int32_t k,
              const float *gout, const int32_t *start, const int32_t *slots,
              float *gfeat) {
  const int64_t wid = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int chunks = (C + 63) >> 6;
  const int64_t m = wid / chunks;
  if (m >= ns) return;  // wave-uniform
  const int c = (int)(wid - m * chunks) * 64 + lane;
  const bool cok = c < C;
  float acc = 0.f;
  const int32_t b = start[m], end = start[m + 1];
  for (int32_t c0 = b; c0 < end; c0 += 64) {
    const int n = end - c0 < 64 ? end - c0 : 64;
    int32_t row = 0;  // grad_out row of the lane's entry
    float we = 0.f;
    if (lane < n) {
      const int32_t e = slots[c0 + lane];
      row = e / k;
      we = w[e];
    }
    int u = 0;
    for (; u + 8 <= n; u += 8) {
      float gv[8], wv[8];
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int64_t r = lane_value(row, u + v);
        wv[v] = lane_value(we, u + v);
        gv[v] = cok ? gout[r * C + c] : 0.f;
      }
#pragma unroll
      for (int v = 0; v < 8; ++v) acc += wv[v] * gv[v];
    }
    for (; u < n; ++u) {
      const int64_t r = lane_value(row, u);
      acc += lane_value(we, u) * (cok ? gout[r * C + c] : 0.f);
    }
  }
  if (cok) gfeat[m * C + c] = acc;
}

// grad_w[i, j] = <grad_out[i, :], feat[idx[i,j], :]>: 16 lanes per entry
// read the two rows coalesced (lane s takes channels s, s + 16, ...) and
// reduce by a fixed butterfly, so the result is reproducible.
constexpr int kDotLanes = 16;

__global__ void k_idw_bw_w(const float *feat, int64_t ns, int32_t C,
                           const int64_t *idx, int64_t nq, int32_t k,
                           const float *gout, float *gw) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t e = t / kDotLanes;
  const int sub = (int)(t % kDotLanes);
  const bool ok = e < nq * k;
  const int64_t m = ok ? idx[e] : -1;
  float acc = 0.f;
  if (m >= 0 && m < ns) {
    const float *f = feat + m * C;
    const float *go = gout + (e / k) * C;
    for (int c = sub; c < C; c += kDotLanes) acc += go[c] * f[c];
  }
  // every lane of the wave takes part (entries past the end add zeros)
#pragma unroll
  for (int o = kDotLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kDotLanes);
  if (ok && sub == 0) gw[e] = acc;
}

// ---------------------------------------------------- fused IDW interpolation
//
// out_i = sum_j w_ij feat[m_ij], w_ij = r_ij / S_i, r_ij = 1 / (d_ij + eps),
// S_i = sum_j r_ij, d_ij = ||p_m - q_i||^2 in the query kernels' fp32 form
// (r = 0 for padded slots).  One wave per query row: lane j owns neighbour j
// (k <= 64), lane c owns channel c of a 64-channel chunk; every sum over j
// or over channels runs in a fixed order (ascending j, or a fixed
// butterfly), so results are reproducible.
constexpr int kRowWaves = 4;

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int kGather = 8;  // neighbour rows in flight per lane (forward)

// v of lane j, j wave-uniform (a scalar read, no LDS round trip)
__device__ __forceinline__ int bitrev4(int v) {
  return ((v & 1) << 3) | ((v & 2) << 1) | ((v & 4) >> 1) | ((v & 8) >> 3);
}

// One halving step of a transposed butterfly: the lanes with bit O clear
// keep (and sum with their partner's) the lower N/2 values, the others the
// upper N/2.
template <int N, int O>
__device__ __forceinline__ void fold(float (&x)[16], int lane) {
  const bool hi = (lane & O) != 0;
#pragma unroll
  for (int t = 0; t < N / 2; ++t) {
    const float send = hi ? x[t] : x[t + N / 2];
    const float keep = hi ? x[t + N / 2] : x[t];
    x[t] = keep + __shfl_xor(send, O, 64);
  }
}

// Sums of x[0..15] over the 64 lanes in 17 shuffles (16 separate wave sums
// take 96), in a fixed order: lane l returns the sum of x[bitrev4(l & 15)].
__device__ __forceinline__ float wave_sum16(float (&x)[16], int lane) {
  fold<16, 1>(x, lane);
  fold<8, 2>(x, lane);
  fold<4, 4>(x, lane);
  fold<2, 8>(x, lane);
  float r = x[0];
  r += __shfl_xor(r, 16, 64);
  r += __shfl_xor(r, 32, 64);
  return r;
}

// Lane j < k: its neighbour m (-1 if padded or out of range) and r_j.
__device__ __forceinline__ void interp_lane(const float *q, int64_t i,
                                            const float *p, int64_t ns,
                                            const int64_t *idx, int32_t k,
                                            float eps, int lane, int &m, float &r) {
  m = -1;
  r = 0.f;
  if (lane < k) {
    const int64_t mm = idx[i * k + lane];
    if (mm >= 0 && mm < ns) {
      m = (int)mm;
      const float dx = __fsub_rn(p[3 * mm], q[3 * i]);
      const float dy = __fsub_rn(p[3 * mm + 1], q[3 * i + 1]);
      const float dz = __fsub_rn(p[3 * mm + 2], q[3 * i + 2]);
      const float d = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)),
                                __fmul_rn(dz, dz));
      r = __fdiv_rn(1.f, __fadd_rn(d, eps));
    }
  }
}

__global__ void __launch_bounds__(64 * kRowWaves)
k_interp_fwd(const float *q, int64_t nq, const float *p, int64_t ns,
             const int64_t *idx, int32_t k, const float *feat, int32_t C,
             float eps, float *w_out, float *out) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * kRowWaves + (threadIdx.x >> 6);
  if (i >= nq) return;  // wave-uniform
  int m;
  float r;
  interp_lane(q, i, p, ns, idx, k, eps, lane, m, r);
  const float S = wave_sum(r);
  const float w = S > 0.f ? __fdiv_rn(r, S) : 0.f;
  if (w_out && lane < k) w_out[i * k + lane] = w;
  for (int c0 = 0; c0 < C; c0 += 64) {
    const int c = c0 + lane;
    float acc = 0.f;
    // kGather neighbour rows in flight; the sum still runs in ascending j
    for (int j0 = 0; j0 < k; j0 += kGather) {
      float v[kGather], wv[kGather];
      bool ok[kGather];
#pragma unroll
      for (int u = 0; u < kGather; ++u) {
        const int j = j0 + u;  // < 64 + kGather: lane j & 63 is only read if j < k
        const int mj = lane_value(m, j & 63);
        wv[u] = __int_as_float(lane_value(__float_as_int(w), j & 63));
        ok[u] = j < k && mj >= 0;
        v[u] = (ok[u] && c < C) ? feat[(int64_t)mj * C + c] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < kGather; ++u)
        if (ok[u]) acc = __fadd_rn(acc, __fmul_rn(wv[u], v[u]));
    }
    if (c < C) out[i * C + c] = acc;
  }
}

// Per row: a_j = <g_i, feat[m_j]>, b = sum_j w_j a_j, dL/dd_j =
// -r_j^2 (a_j - b) / S, stored to gd (and w to w_out) for the support-side
// sums; dL/dq_i = sum_j 2 dL/dd_j (q_i - p_j).
__global__ void __launch_bounds__(64 * kRowWaves)
k_interp_bw_row(const float *q, int64_t nq, const float *p, int64_t ns,
                const int64_t *idx, int32_t k, const float *feat, int32_t C,
                float eps, const float *gout, float *w_out, float *gd,
                float *gq) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * kRowWaves + (threadIdx.x >> 6);
  if (i >= nq) return;  // wave-uniform
  int m;
  float r;
  interp_lane(q, i, p, ns, idx, k, eps, lane, m, r);
  const float S = wave_sum(r);
  const float w = S > 0.f ? __fdiv_rn(r, S) : 0.f;
  // a_j for 16 neighbours at a time: the 16 gathers in flight, then one
  // transposed butterfly for the 16 dot products
  float a = 0.f;
  for (int j0 = 0; j0 < k; j0 += 16) {
    float x[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) x[u] = 0.f;
    for (int c0 = 0; c0 < C; c0 += 64) {
      const int c = c0 + lane;
      const float go = c < C ? gout[i * C + c] : 0.f;
      float v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int j = j0 + u;  // < 80: lane j & 63 is only read if j < k
        const int mj = lane_value(m, j & 63);
        v[u] = (j < k && mj >= 0 && c < C) ? feat[(int64_t)mj * C + c] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) x[u] = __fadd_rn(x[u], __fmul_rn(go, v[u]));
    }
    const float red = wave_sum16(x, lane);
    const float aj = __shfl(red, bitrev4(lane & 15), 64);
    if (lane >= j0 && lane < j0 + 16) a = aj;
  }
  const float b = wave_sum(lane < k ? __fmul_rn(w, a) : 0.f);
  float g = 0.f;
  if (m >= 0 && S > 0.f) g = -__fmul_rn(__fmul_rn(r, r), __fdiv_rn(__fsub_rn(a, b), S));
  if (lane < k) {
    if (w_out) w_out[i * k + lane] = w;
    if (gd) gd[i * k + lane] = g;
  }
  if (gq) {
    float gx = 0.f, gy = 0.f, gz = 0.f;
    if (m >= 0) {
      const float c2 = __fmul_rn(2.f, g);
      gx = __fmul_rn(c2, __fsub_rn(q[3 * i], p[3 * (int64_t)m]));
      gy = __fmul_rn(c2, __fsub_rn(q[3 * i + 1], p[3 * (int64_t)m + 1]));
      gz = __fmul_rn(c2, __fsub_rn(q[3 * i + 2], p[3 * (int64_t)m + 2]));
    }
    gx = wave_sum(gx);
    gy = wave_sum(gy);
    gz = wave_sum(gz);
    if (lane == 0) {
      gq[3 * i] = gx;
      gq[3 * i + 1] = gy;
      gq[3 * i + 2] = gz;
    }
  }
}

// Reverse CSR of idx: segment m = [start[m], start[m+1]) of *sorted holds the
// entries e with idx[e] == m, ascending.
hipError_t build_reverse_csr(const int64_t *idx, int64_t E, int64_t ns,
                             const Workspace &w, const int32_t **sorted,
                             hipStream_t s) {
  *sorted = w.rank;
  hipLaunchKernelGGL(k_zero, dim3(nblocks(ns) > 4096 ? 4096 : nblocks(ns)),
                     dim3(kBlock), 0, s, w.counts, ns);
  if (E > 0)
    hipLaunchKernelGGL(k_rc_count, dim3(nblocks(E)), dim3(kBlock), 0, s, idx, E,
                       ns, w.counts, w.rank);
  hipError_t err = launch_scan_host(w.counts, w.start, ns, w.part, s);
  if (err != hipSuccess) return err;
  if (E > 0)
    hipLaunchKernelGGL(k_rc_place, dim3(nblocks(E)), dim3(kBlock), 0, s, idx, E,
                       w.start, w.rank, w.slots);
  // sorted segments go to w.rank (free once placed); w.slots then points there
  if (ns > 0 && E > 0) {
    const int64_t nb = (ns + kRankWaves - 1) / kRankWaves;
    hipLaunchKernelGGL(k_rc_rank, dim3((unsigned)nb), dim3(64 * kRankWaves), 0, s,
                       w.start, ns, w.slots, w.rank);
  }
  return hipGetLastError();
}

}  // namespace

int64_t backward_workspace_bytes(int64_t nq, int64_t ns, int32_t k) {
  const int64_t E = nq * k;
  return align_up(4 * ns) + align_up(4 * (ns + 1)) +
         align_up(4 * scan_parts_needed(ns)) + align_up(4 * E) +
         align_up(4 * E) + 256;
}

hipError_t launch_dist_backward(const float *q, int64_t nq, const float *p,
                                int64_t ns, const int64_t *idx, int32_t k,
                                const float *grad_d, float *grad_q,
                                float *grad_p, void *ws, hipStream_t s) {
  const int64_t E = nq * k;
  if (grad_q && nq > 0)
    hipLaunchKernelGGL(k_dist_bw_q, dim3(nblocks(nq)), dim3(kBlock), 0, s, q, nq,
                       p, ns, idx, k, grad_d, grad_q);
  if (grad_p && ns > 0) {
    const Workspace w = carve(ws, ns, E);
    const int32_t *sorted = nullptr;
    hipError_t err = build_reverse_csr(idx, E, ns, w, &sorted, s);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_dist_bw_p, dim3(nblocks(ns * 64)), dim3(kBlock), 0, s, q, p,
                       ns, k, grad_d, w.start, sorted, grad_p);
  }
  return hipGetLastError();
}

hipError_t launch_idw_forward(const float *feat, int64_t ns, int32_t C,
                              const int64_t *idx, const float *w, int64_t nq,
                              int32_t k, float *out, hipStream_t s) {
  (void)ns;
  if (nq * C > 0)
    hipLaunchKernelGGL(k_idw_fwd, dim3(nblocks(nq * C)), dim3(kBlock), 0, s,
                       feat, ns, C, idx, w, nq, k, out);
  return hipGetLastError();
}

hipError_t launch_idw_backward(const float *feat, int64_t ns, int32_t C,
                               const int64_t *idx, const float *w, int64_t nq,
                               int32_t k, const float *grad_out,
                               float *grad_feat, float *grad_w, void *ws,
                               hipStream_t s) {
  const int64_t E = nq * k;
  if (grad_w && E > 0)
    hipLaunchKernelGGL(k_idw_bw_w, dim3(nblocks(E * kDotLanes)), dim3(kBlock), 0, s,
                       feat, ns, C, idx, nq, k, grad_out, grad_w);
  if (grad_feat && ns * C > 0) {
    const Workspace wsp = carve(ws, ns, E);
    const int32_t *sorted = nullptr;
    hipError_t err = build_reverse_csr(idx, E, ns, wsp, &sorted, s);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_idw_bw_feat, dim3(nblocks(ns * ((C + 63) / 64) * 64)), dim3(kBlock), 0, s,
                       ns, C, w, k, grad_out, wsp.start, sorted, grad_feat);
  }
  return hipGetLastError();
}

int64_t interp_workspace_bytes(int64_t nq, int64_t ns, int32_t k) {
  const int64_t E = nq * k;
  return 2 * align_up(4 * E) + backward_workspace_bytes(nq, ns, k);
}

hipError_t launch_interp_forward(const float *q, int64_t nq, const float *p,
                                 int64_t ns, const int64_t *idx, int32_t k,
                                 const float *feat, int32_t C, float eps,
                                 float *w_out, float *out, hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  const int64_t nb = (nq + kRowWaves - 1) / kRowWaves;
  hipLaunchKernelGGL(k_interp_fwd, dim3((unsigned)nb), dim3(64 * kRowWaves), 0, s,
                     q, nq, p, ns, idx, k, feat, C, eps, w_out, out);
  return hipGetLastError();
}

hipError_t launch_interp_backward(const float *q, int64_t nq, const float *p,
                                  int64_t ns, const int64_t *idx, int32_t k,
                                  const float *feat, int32_t C, float eps,
                                  const float *grad_out, float *grad_feat,
                                  float *grad_q, float *grad_p, void *ws,
                                  hipStream_t s) {
  const int64_t E = nq * k;
  char *base = (char *)ws;
  float *w = (float *)base;
  float *gd = (float *)(base + align_up(4 * E));
  void *csr_ws = base + 2 * align_up(4 * E);
  if (nq > 0) {
    const int64_t nb = (nq + kRowWaves - 1) / kRowWaves;
    hipLaunchKernelGGL(k_interp_bw_row, dim3((unsigned)nb), dim3(64 * kRowWaves), 0, s,
                       q, nq, p, ns, idx, k, feat, C, eps, grad_out, w, gd, grad_q);
  }
  if ((grad_p || grad_feat) && ns > 0) {
    const Workspace wsp = carve(csr_ws, ns, E);
    const int32_t *sorted = nullptr;
    hipError_t err = build_reverse_csr(idx, E, ns, wsp, &sorted, s);
    if (err != hipSuccess) return err;
    if (grad_p)
      hipLaunchKernelGGL(k_dist_bw_p, dim3(nblocks(ns * 64)), dim3(kBlock), 0, s, q, p, ns,
                         k, gd, wsp.start, sorted, grad_p);
    if (grad_feat && ns * C > 0)
      hipLaunchKernelGGL(k_idw_bw_feat, dim3(nblocks(ns * ((C + 63) / 64) * 64)), dim3(kBlock), 0, s,
                         ns, C, w, k, grad_out, wsp.start, sorted, grad_feat);
  }
  return hipGetLastError();
}

}  // namespace fk
