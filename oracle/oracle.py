"""CPU oracle for the FlashKNN grid-kNN path (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.  The shipped
package never imports it; its kernels must run on the GPU or fail.

Two independent layers:

* ``fk_oracle.c`` (built by ``oracle/Makefile`` with -ffp-contract=off) holds
  the fast restatements: Morton keys, the 27-cell candidate-set kNN of
  FlashKNN.query and the exact per-segment kNN of pointops.knn_query.
* this module adds slow, independent restatements used to *pin* the C code:
  a bit-loop Morton key, a pure-Python candidate enumeration for tiny
  inputs and a vectorised numpy fp32 brute force; plus the closed forms of
  the distance backward and of inverse-distance-weighted interpolation.

Reference citations:
  Morton key          FlashKNN/csrc/flash_knn_construction.cu:13-27,
                      FlashKNN/functions/z_order.py:32-42,57-91
  neighbour keys      FlashKNN/csrc/flash_knn_construction.cu:42-63
  candidate set       FlashKNN/csrc/flash_knn_query_dynamic_load.cu:87-155,
                      469-652; FlashKNN/functions/FlashKnnWrapper.py:226-337
  distance            dynamic_load.cu:601-604 (dx = p - q, ((x+y)+z))
  strict cut          dynamic_load.cu:606
  padding (self,inf)  dynamic_load.cu:562; FlashKnnWrapper.py:427-429
  selected query      FlashKnnWrapper.py:646-849
  back query          FlashKnnWrapper.py:460-644 (cells = trunc(xyz/size))
  voxel-first         Networks/hierarchy.py:23-36 (first index per voxel key)
  pointops knn        Pointcept/libs/pointops/src/knn_query/
                      knn_query_cuda_kernel.cu:60-104; functions/query.py:7-24
  IDW                 Pointcept/libs/pointops/functions/interpolation.py:8-56
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfk_oracle.so")
_lib = None

_f32p = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build():
    """Compile fk_oracle.c into oracle/_build (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.fko_morton_encode.argtypes = [_i64p, _i64p, ctypes.c_int64, _i64p]
        L.fko_morton_neighbor_keys.argtypes = [_i64p, _i64p, ctypes.c_int64, _i64p]
        L.fko_grid_knn.argtypes = [
            ctypes.c_int64, _f32p, _i64p, _i64p, ctypes.c_int64, _f32p, _i64p,
            _i64p, _i64p, ctypes.c_int, ctypes.c_int, ctypes.c_float, _i64p, _f32p]
        L.fko_grid_knn.restype = ctypes.c_int
        L.fko_brute_knn.argtypes = [
            ctypes.c_int64, _f32p, _i64p, ctypes.c_int64, _f32p, _i64p,
            ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int64,
            ctypes.c_int, _i64p, ctypes.c_void_p]
        L.fko_brute_knn.restype = ctypes.c_int
        L.fko_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def num_threads() -> int:
    return int(lib().fko_num_threads())


# ------------------------------------------------------------------ Morton

def morton_encode(grid, batch=None):
    """48-bit Morton key of int grid coords plus batch << 48 (C restatement)."""
    g = _i64(grid).reshape(-1, 3)
    b = None if batch is None else _i64(batch).reshape(-1)
    out = np.empty(len(g), np.int64)
    lib().fko_morton_encode(_p(g, _i64p), _p(b, _i64p), len(g), _p(out, _i64p))
    return out


def morton_neighbor_keys(grid_down, parent_batch=None):
    g = _i64(grid_down).reshape(-1, 3)
    b = None if parent_batch is None else _i64(parent_batch).reshape(-1)
    out = np.empty(len(g) * 27, np.int64)
    lib().fko_morton_neighbor_keys(_p(g, _i64p), _p(b, _i64p), len(g), _p(out, _i64p))
    return out


def morton_key_bitloop(x: int, y: int, z: int, b: int = 0) -> int:
    """Independent textbook interleave: bit i of x/y/z -> bit 3i+2/3i+1/3i."""
    key = 0
    for i in range(16):
        key |= ((x >> i) & 1) << (3 * i + 2)
        key |= ((y >> i) & 1) << (3 * i + 1)
        key |= ((z >> i) & 1) << (3 * i)
    return key | (b << 48)


# ------------------------------------------------------- grid candidate kNN

def grid_knn(sxyz, scell, sbatch, qxyz, qcell, qbatch, qpad, k, R=1,
             cut_r2=float("inf")):
    """K smallest (d, support index) among the support points of the same
    batch whose cell is within R of the query's cell on every axis."""
    sxyz, scell = _f32(sxyz).reshape(-1, 3), _i64(scell).reshape(-1, 3)
    qxyz, qcell = _f32(qxyz).reshape(-1, 3), _i64(qcell).reshape(-1, 3)
    sb = None if sbatch is None else _i64(sbatch).reshape(-1)
    qb = None if qbatch is None else _i64(qbatch).reshape(-1)
    pad = _i64(qpad).reshape(-1)
    nq = len(qxyz)
    idx = np.empty((nq, k), np.int64)
    d = np.empty((nq, k), np.float32)
    lib().fko_grid_knn(len(sxyz), _p(sxyz, _f32p), _p(scell, _i64p), _p(sb, _i64p),
                       nq, _p(qxyz, _f32p), _p(qcell, _i64p), _p(qb, _i64p),
                       _p(pad, _i64p), int(R), int(k), float(cut_r2),
                       _p(idx, _i64p), _p(d, _f32p))
    return idx, d


def flashknn_query(grid_coord, batch_idx, xyz, k, num_down, cut_radius=float("inf")):
    """FlashKNN.query (FlashKnnWrapper.py:165-449): self-query, search cell =
    grid_coord >> num_down, padding = the query's own index.  When ``xyz`` is
    None the distances are taken on grid_coord as float (:353-354)."""
    g = _i64(grid_coord).reshape(-1, 3)
    cell = g >> int(num_down)
    pts = g.astype(np.float32) if xyz is None else _f32(xyz)
    n = len(g)
    pad = np.arange(n, dtype=np.int64)
    return grid_knn(pts, cell, batch_idx, pts, cell, batch_idx, pad, k, 1,
                    cut_radius * cut_radius)


def selected_query(xyz, grid_coord, query_indices, batch_idx, k, num_down,
                   cut_radius=float("inf")):
    """FlashKNN.selected_query (FlashKnnWrapper.py:646-849): queries are the
    support points listed in query_indices; padding = the query's own index."""
    g = _i64(grid_coord).reshape(-1, 3)
    cell = g >> int(num_down)
    pts = g.astype(np.float32) if xyz is None else _f32(xyz)
    qi = _i64(query_indices).reshape(-1)
    qb = None if batch_idx is None else _i64(batch_idx)[qi]
    return grid_knn(pts, cell, batch_idx, pts[qi], cell[qi], qb, qi, k, 1,
                    cut_radius * cut_radius)


def back_query(xyz, sel_indices, query_grid_size, batch_idx, k):
    """FlashKNN.back_query (FlashKnnWrapper.py:460-644): every point of xyz
    queries the subset xyz[sel_indices]; cells are (xyz/query_grid_size).long()
    (:480): trunc of the correctly rounded fp32 quotient; returned indices
    are positions in sel_indices.  Padding (-1, +inf)
    (dynamic_load.cu:1098-1102)."""
    pts = _f32(xyz).reshape(-1, 3)
    cell = np.trunc(pts / np.float32(query_grid_size)).astype(np.int64)
    sel = _i64(sel_indices).reshape(-1)
    sb = None if batch_idx is None else _i64(batch_idx)[sel]
    pad = np.full(len(pts), -1, np.int64)
    return grid_knn(pts[sel], cell[sel], sb, pts, cell, batch_idx, pad, k, 1)


# ------------------------------------------------- voxel-first subsampling

def voxel_keys(xyz, size, offsets=None):
    """Voxel key of every point: (cloud, low 16 bits of floor(x / size) per
    axis, fp32 quotient correctly rounded as torch.floor(xyz / size) divides).  The reference's key is xyz2key of the voxel coordinates
    (FlashKNN/functions/z_order.py:56-91, depth 16: bits 0..15 of each axis,
    a bijection of them), so equal keys here <=> equal keys there."""
    pts = _f32(xyz).reshape(-1, 3)
    v = np.floor(pts / np.float32(size)).astype(np.int64) & 0xFFFF
    cloud = np.zeros(len(pts), np.int64)
    if offsets is not None:
        cloud = np.searchsorted(_i64(offsets).reshape(-1), np.arange(len(pts)), side="right")
    return (cloud << 48) | (v[:, 0] << 32) | (v[:, 1] << 16) | v[:, 2]


def voxel_first(xyz, size, offsets=None):
    """_voxel_first_cuda (Networks/hierarchy.py:23-36): a stable argsort of
    the voxel keys, the first of every run (the smallest index of each
    voxel), returned ascending.  np.unique's return_index is that first
    occurrence."""
    key = voxel_keys(xyz, size, offsets)
    _, first = np.unique(key, return_index=True)
    return np.sort(first).astype(np.int64)


def voxel_first_python(xyz, size, offsets=None):
    """Pure-Python loop form of voxel_first (small inputs): scan the points in
    index order and keep each voxel's first."""
    key = voxel_keys(xyz, size, offsets)
    seen, keep = set(), []
    for i, k in enumerate(key.tolist()):
        if k not in seen:
            seen.add(k)
            keep.append(i)
    return np.array(keep, np.int64)


# ---------------------------------------------------- exact brute force kNN

def brute_knn(sxyz, soff, qxyz, qoff, k, cut_r2=float("inf"), pad=-1, f64=False):
    """Exact per-segment kNN (pointops offsets = cumulative segment ends)."""
    sxyz, qxyz = _f32(sxyz).reshape(-1, 3), _f32(qxyz).reshape(-1, 3)
    so, qo = _i64(soff).reshape(-1), _i64(qoff).reshape(-1)
    assert len(so) == len(qo)
    nq = len(qxyz)
    idx = np.empty((nq, k), np.int64)
    d = np.empty((nq, k), np.float64 if f64 else np.float32)
    lib().fko_brute_knn(len(sxyz), _p(sxyz, _f32p), _p(so, _i64p), nq,
                        _p(qxyz, _f32p), _p(qo, _i64p), len(so), int(k),
                        float(cut_r2), int(pad), int(bool(f64)), _p(idx, _i64p),
                        d.ctypes.data_as(ctypes.c_void_p))
    return idx, d


def brute_knn_numpy(sxyz, qxyz, k):
    """Independent numpy fp32 restatement for one segment: the same
    ((dx*dx + dy*dy) + dz*dz) evaluated elementwise in float32 (numpy never
    contracts), ranked by lexsort on (d, index)."""
    s, q = _f32(sxyz).reshape(-1, 3), _f32(qxyz).reshape(-1, 3)
    dx = s[None, :, 0] - q[:, None, 0]
    dy = s[None, :, 1] - q[:, None, 1]
    dz = s[None, :, 2] - q[:, None, 2]
    d = (dx * dx + dy * dy) + dz * dz
    assert d.dtype == np.float32
    n = s.shape[0]
    out_i = np.full((len(q), k), -1, np.int64)
    out_d = np.full((len(q), k), np.inf, np.float32)
    ar = np.arange(n)
    for r in range(len(q)):
        o = np.lexsort((ar, d[r]))[:k]
        out_i[r, :len(o)] = o
        out_d[r, :len(o)] = d[r, o]
    return out_i, out_d


def grid_knn_python(sxyz, scell, sbatch, qxyz, qcell, qbatch, qpad, k, R=1,
                    cut_r2=float("inf")):
    """Pure-Python candidate enumeration (tiny inputs only): test every
    support point for same batch and |cell delta| <= R, fp32 distance via
    numpy scalars, sort (d, index)."""
    f = np.float32
    out_i = np.empty((len(qxyz), k), np.int64)
    out_d = np.empty((len(qxyz), k), np.float32)
    for q in range(len(qxyz)):
        cand = []
        for s in range(len(sxyz)):
            if sbatch is not None and int(sbatch[s]) != int(qbatch[q]):
                continue
            if max(abs(int(scell[s][a]) - int(qcell[q][a])) for a in range(3)) > R:
                continue
            dx = f(sxyz[s][0]) - f(qxyz[q][0])
            dy = f(sxyz[s][1]) - f(qxyz[q][1])
            dz = f(sxyz[s][2]) - f(qxyz[q][2])
            d = f(f(f(dx * dx) + f(dy * dy)) + f(dz * dz))
            if d < f(cut_r2):
                cand.append((float(d), s))
        cand.sort()
        for j in range(k):
            if j < len(cand):
                out_d[q, j], out_i[q, j] = cand[j]
            else:
                out_d[q, j], out_i[q, j] = np.inf, qpad[q]
    return out_i, out_d


# ------------------------------------------------------- backward closed forms

def dist_backward(q, p, idx, grad, valid=None):
    """Closed form of d_ij = ||q_i - p_idx(i,j)||^2 (float64):
    dL/dq_i = sum_j 2 g_ij (q_i - p_j);  dL/dp_m = -sum_{idx=m} 2 g_ij (q_i - p_m).
    Entries with idx < 0 (or valid == False) contribute nothing."""
    q = np.asarray(q, np.float64)
    p = np.asarray(p, np.float64)
    g = np.asarray(grad, np.float64)
    idx = np.asarray(idx, np.int64)
    ok = idx >= 0 if valid is None else (np.asarray(valid) & (idx >= 0))
    safe = np.where(ok, idx, 0)
    diff = q[:, None, :] - p[safe]                      # [nq, K, 3]
    c = 2.0 * g * ok                                    # [nq, K]
    gq = (c[..., None] * diff).sum(1)
    gp = np.zeros_like(p)
    np.add.at(gp, safe.reshape(-1), -(c[..., None] * diff).reshape(-1, 3))
    return gq, gp


def idw_weights(d2, eps):
    """Smooth inverse-distance weights on squared distance: w = 1/(d2 + eps),
    normalised per row (differentiable at d2 = 0)."""
    w = 1.0 / (np.asarray(d2, np.float64) + eps)
    return w / w.sum(1, keepdims=True)


def idw_forward(feat, idx, d2, eps):
    """out_i = sum_j wn_ij feat[idx_ij] with wn = idw_weights(d2)."""
    wn = idw_weights(d2, eps)
    f = np.asarray(feat, np.float64)
    return (wn[..., None] * f[np.asarray(idx)]).sum(1)


def pointops_interpolation_weights(dist):
    """pointops interpolation weights on sqrt distance: 1/(dist + 1e-8),
    normalised (Pointcept/libs/pointops/functions/interpolation.py:15-17)."""
    r = 1.0 / (np.asarray(dist, np.float64) + 1e-8)
    return r / r.sum(1, keepdims=True)


# ------------------------------------------- adaptive per-point levels

def point_candidates(cell, batch=None, R=1):
    """|C(q)| for every point: points of the same batch whose cell is within
    R per axis (numpy: unique cells, searchsorted of the (2R+1)^3 offsets)."""
    c = _i64(cell).reshape(-1, 3)
    b = np.zeros(len(c), np.int64) if batch is None else _i64(batch).reshape(-1)
    c = c - c.min(0) + R
    W = c.max(0) + R + 1
    key = ((b * W[0] + c[:, 0]) * W[1] + c[:, 1]) * W[2] + c[:, 2]
    uk, cnt = np.unique(key, return_counts=True)
    total = np.zeros(len(c), np.int64)
    for dx in range(-R, R + 1):
        for dy in range(-R, R + 1):
            for dz in range(-R, R + 1):
                nk = ((b * W[0] + c[:, 0] + dx) * W[1] + c[:, 1] + dy) * W[2] + c[:, 2] + dz
                pos = np.searchsorted(uk, nk)
                ok = pos < len(uk)
                hit = np.zeros(len(c), bool)
                hit[ok] = uk[pos[ok]] == nk[ok]
                total += np.where(hit, cnt[np.minimum(pos, len(uk) - 1)], 0)
    return total


def adaptive_query(grid_coord, batch_idx, xyz, k, min_f=2, max_f=8, max_shift=None,
                   cut_radius=float("inf")):
    """AdaptiveNeighborhoodFlashKNN.query
    (FlashKNN/functions/adaptive_octree.py:129-146 shift per batch,
    :204-208 27-node counts, :312-351 level selection, then the 27-node
    query at the selected level).  Returns (idx, d, selected level)."""
    g = _i64(grid_coord).reshape(-1, 3)
    b = np.zeros(len(g), np.int64) if batch_idx is None else _i64(batch_idx)
    B = int(b.max()) + 1
    mins = np.full((B, 3), np.iinfo(np.int64).max, np.int64)
    np.minimum.at(mins, b, g)
    g = g - mins[b]
    if max_shift is None:
        max_shift = int(g.max()).bit_length()
    counts = [point_candidates(g >> s, b) for s in range(max_shift + 1)]
    sel = np.full(len(g), max_shift, np.int64)
    active = counts[max_shift] > max_f * k
    for s in range(max_shift - 1, -1, -1):
        use = active & (counts[s] >= min_f * k)
        sel[use] = s
        active = use & (counts[s] > max_f * k)
    pts = g.astype(np.float32) if xyz is None else _f32(xyz)
    idx = np.empty((len(g), k), np.int64)
    d = np.empty((len(g), k), np.float32)
    for s in np.unique(sel):
        q = np.where(sel == s)[0]
        cell = g >> s
        qi, qd = grid_knn(pts, cell, b, pts[q], cell[q], b[q], q, k, 1,
                          cut_radius * cut_radius)
        idx[q], d[q] = qi, qd
    return idx, d, sel
