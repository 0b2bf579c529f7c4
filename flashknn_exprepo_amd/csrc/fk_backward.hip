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

inline unsigned nblocks(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  return (unsigned)(b < 1 ? 1 : b);
}

struct Workspace {
  int32_t *counts;  // [ns]
  int32_t *start;   // [ns + 1]
  int32_t *part;    // scan partials
  int32_t *rank;    // [E]
  int32_t *slots;   // [E]
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

// Insertion sort of each segment by entry id: fixes the summation order.
__global__ void k_rc_sort(const int32_t *start, int64_t ns, int32_t *slots) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= ns) return;
  const int32_t b = start[m], e = start[m + 1];
  for (int32_t i = b + 1; i < e; ++i) {
    const int32_t v = slots[i];
    int32_t j = i - 1;
    while (j >= b && slots[j] > v) {
      slots[j + 1] = slots[j];
      --j;
    }
    slots[j + 1] = v;
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

__global__ void k_dist_bw_p(const float *q, const float *p, int64_t ns,
                            int32_t k, const float *g, const int32_t *start,
                            const int32_t *slots, float *gp) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= ns) return;
  const float px = p[3 * m], py = p[3 * m + 1], pz = p[3 * m + 2];
  float ax = 0.f, ay = 0.f, az = 0.f;
  for (int32_t t = start[m]; t < start[m + 1]; ++t) {
    const int32_t e = slots[t];
    const int64_t i = e / k;
    const float c = 2.f * g[e];
    ax -= c * (q[3 * i] - px);
    ay -= c * (q[3 * i + 1] - py);
    az -= c * (q[3 * i + 2] - pz);
  }
  gp[3 * m] = ax;
  gp[3 * m + 1] = ay;
  gp[3 * m + 2] = az;
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
__global__ void k_idw_bw_feat(int64_t ns, int32_t C, const float *w, int32_t k,
                              const float *gout, const int32_t *start,
                              const int32_t *slots, float *gfeat) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ns * C) return;
  const int64_t m = t / C;
  const int c = (int)(t - m * C);
  float acc = 0.f;
  for (int32_t s = start[m]; s < start[m + 1]; ++s) {
    const int32_t e = slots[s];
    acc += w[e] * gout[(int64_t)(e / k) * C + c];
  }
  gfeat[t] = acc;
}

// grad_w[i, j] = <grad_out[i, :], feat[idx[i,j], :]>  (thread per entry).
__global__ void k_idw_bw_w(const float *feat, int64_t ns, int32_t C,
                           const int64_t *idx, int64_t nq, int32_t k,
                           const float *gout, float *gw) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nq * k) return;
  const int64_t m = idx[e];
  float acc = 0.f;
  if (m >= 0 && m < ns) {
    const float *f = feat + m * C;
    const float *go = gout + (e / k) * C;
    for (int c = 0; c < C; ++c) acc += go[c] * f[c];
  }
  gw[e] = acc;
}

hipError_t build_reverse_csr(const int64_t *idx, int64_t E, int64_t ns,
                             const Workspace &w, hipStream_t s) {
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
  if (ns > 0)
    hipLaunchKernelGGL(k_rc_sort, dim3(nblocks(ns)), dim3(kBlock), 0, s,
                       w.start, ns, w.slots);
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
    hipError_t err = build_reverse_csr(idx, E, ns, w, s);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_dist_bw_p, dim3(nblocks(ns)), dim3(kBlock), 0, s, q, p,
                       ns, k, grad_d, w.start, w.slots, grad_p);
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
    hipLaunchKernelGGL(k_idw_bw_w, dim3(nblocks(E)), dim3(kBlock), 0, s, feat, ns,
                       C, idx, nq, k, grad_out, grad_w);
  if (grad_feat && ns * C > 0) {
    const Workspace wsp = carve(ws, ns, E);
    hipError_t err = build_reverse_csr(idx, E, ns, wsp, s);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_idw_bw_feat, dim3(nblocks(ns * C)), dim3(kBlock), 0, s,
                       ns, C, w, k, grad_out, wsp.start, wsp.slots, grad_feat);
  }
  return hipGetLastError();
}

}  // namespace fk
