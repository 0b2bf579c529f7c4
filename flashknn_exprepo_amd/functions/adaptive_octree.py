"""Adaptive (per-point octree level) neighbourhoods
(drop-in for FlashKNN/functions/adaptive_octree.py:211-612).

Semantics kept from the reference:

* coordinates are shifted per batch to be non-negative (:129-146); level s
  groups points by ``coord >> s`` within a batch, s = 0 .. max_shift
  (default: the bit length of the largest shifted coordinate, :244-246);
* the statistic of a node is its 27-node candidate count (:204-208);
* level selection (:312-351): start at the coarsest level; a point keeps
  descending while its node's count exceeds ``max_candidates_factor * k``,
  and accepts a child level only if the child's count is at least
  ``min_candidates_factor * k``;
* the query of a point scans the 27 nodes around its node at its selected
  level (the reference flattens the levels into one launch, :369-488).

Build: instead of a Morton octree with neighbour propagation, every level
is one dense-box grid of libflashknn (cells = shifted coord >> s), the
per-point counts come from fk_point_candidates, and each selected level is
one attached-query launch on its grid.  Results are the K nearest of each
point's candidate set with ties to the smaller index.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Any

import numpy as np
import torch

from .. import _lib as F
from ..grid import Grid

_MAX_COORD = 1 << 16
_MAX_BATCH = 32768
_RESERVE_CAP = 1 << 28   # largest dense box a level reserves (cells)
_SYNC_CELLS = 1 << 21    # boxes above this are built only if some point needs them
_COLUMN_CELLS_PER_POINT = 64  # separately built levels beyond this take the column layout
_CHUNK = 1024   # points per chunk of the per-batch minimum
_SPARE = 4096   # spare rows that absorb the atomics of unused lanes


def walk_levels(c: torch.Tensor, shifts: torch.Tensor, lo: int, hi: int):
    """The coarse-to-fine level walk over consecutive levels at once.

    ``c`` [L, n]: candidate counts at levels shifts[0] > shifts[1] > ...
    (coarse first).  Returns (sel int16 [n], active bool [n]) equal to the
    walk ``active = c[0] > hi``, then per level i >= 1
    ``use = active & (c[i] >= lo); sel[use] = shifts[i];
    active = use & (c[i] > hi)``, starting from sel = shifts[0]: a point is
    active after level i iff c[0] > hi and c[j] > hi, c[j] >= lo for
    1 <= j <= i (a prefix AND, i.e. a running minimum), and it takes the last
    level i >= 1 where it was active before and reached lo."""
    ge = c >= lo
    keep = c > hi
    keep[1:] &= ge[1:]
    run = keep.to(torch.int32).cummin(0).values
    sh = shifts.to(torch.int16)
    if c.shape[0] > 1:
        use = run[:-1].bool() & ge[1:]
        rank = torch.arange(1, c.shape[0], device=c.device, dtype=torch.int32)[:, None]
        last = (use.to(torch.int32) * rank).amax(0)
        sel = sh[last.long()]
    else:
        sel = sh[0].expand(c.shape[1]).clone()
    return sel, run[-1].bool()


def _batch_reduce(v: torch.Tensor, batch: torch.Tensor, B: int, op: str) -> torch.Tensor:
    n = v.shape[0]
    dev = v.device
    nc = (n + _CHUNK - 1) // _CHUNK
    pad = nc * _CHUNK - n
    vp = torch.cat([v, v[-1:].expand(pad, 3)]) if pad else v
    vc = vp.view(nc, _CHUNK, 3)
    red = vc.amin(1) if op == "amin" else vc.amax(1)
    if B == 1:
        # two stages: torch reduces the rows of an [N, 3] tensor on a few
        # blocks only (~50 us for 92k points)
        return red.amin(0, keepdim=True) if op == "amin" else red.amax(0, keepdim=True)
    bp = torch.cat([batch, batch[-1:].expand(pad)]) if pad else batch
    bc = bp.view(nc, _CHUNK)
    blo = bc.amin(1)
    uni = blo == bc.amax(1)
    info = torch.iinfo(torch.int64)
    out = torch.full((B + _SPARE, 3), info.max if op == "amin" else info.min,
                     dtype=torch.int64, device=dev)
    spare_c = B + torch.arange(nc, device=dev) % _SPARE
    idx_c = torch.where(uni, blo, spare_c)
    out.scatter_reduce_(0, idx_c[:, None].expand(-1, 3), red, op, include_self=True)
    spare_p = B + torch.arange(n, device=dev) % _SPARE
    uni_p = uni[:, None].expand(nc, _CHUNK).reshape(-1)[:n]
    idx_p = torch.where(uni_p, spare_p, batch)
    out.scatter_reduce_(0, idx_p[:, None].expand(-1, 3), v, op, include_self=True)
    return out[:B].contiguous()


def batch_min(v: torch.Tensor, batch: torch.Tensor, B: int) -> torch.Tensor:
    """Per-batch minimum of v (int64 [N, 3]) -> int64 [B, 3], no host sync.

    A scatter-min of every point onto B rows serialises N atomics on a few
    addresses.  Here a chunk of consecutive points with one batch index
    contributes its minimum once; points of chunks that mix batches are
    scattered one by one, and every other scatter lane lands in one of
    ``_SPARE`` rows that are dropped, so no row sees more than a few
    hundred atomics when batches are stored contiguously."""
    return _batch_reduce(v, batch, B, "amin")


def batch_max(v: torch.Tensor, batch: torch.Tensor, B: int) -> torch.Tensor:
    """Per-batch maximum of v (int64 [N, 3]) -> int64 [B, 3], as batch_min."""
    return _batch_reduce(v, batch, B, "amax")


@dataclass
class AdaptiveOctreeHierarchy:
    grid: torch.Tensor               # shifted coordinates, int64 [N, 3]
    batch: torch.Tensor              # int64 [N]
    num_batches: int
    coordinate_offsets: torch.Tensor  # per-batch minimum, int64 [B, 3]
    max_shift: int
    candidate_counts: list = field(default_factory=list)  # int32 [N] per level
    grids: list = field(default_factory=list)             # Grid per level
    construction_ms: float = 0.0

    @property
    def levels(self):
        return list(range(self.max_shift + 1))


def _validate(grid_coord, batch_idx):
    if not grid_coord.is_cuda or not batch_idx.is_cuda:
        raise ValueError("adaptive octree inputs must be HIP device tensors")
    if grid_coord.dtype != torch.int64 or batch_idx.dtype != torch.int64:
        raise ValueError("grid_coord and batch_idx must use torch.int64")
    if grid_coord.ndim != 2 or grid_coord.shape[1] != 3:
        raise ValueError("grid_coord must have shape [N, 3]")
    if batch_idx.ndim != 1 or batch_idx.shape[0] != grid_coord.shape[0]:
        raise ValueError("batch_idx must have shape [N]")
    if grid_coord.shape[0] == 0:
        raise ValueError("cannot construct an octree for an empty point cloud")
    if bool((batch_idx < 0).any()) or bool((batch_idx >= _MAX_BATCH).any()):
        raise ValueError("batch indices must lie in [0, 32768)")


@torch.no_grad()
def build_adaptive_octree(grid_coord: torch.Tensor, batch_idx: torch.Tensor,
                          max_shift: int | None = None, xyz: torch.Tensor | None = None,
                          ) -> AdaptiveOctreeHierarchy:
    """One grid per level plus the per-point 27-node candidate counts."""
    _validate(grid_coord, batch_idx)
    torch.cuda.synchronize(grid_coord.device)
    t0 = time.perf_counter()
    B = int(batch_idx.max()) + 1
    mins = batch_min(grid_coord, batch_idx, B)
    g = (grid_coord - mins[batch_idx]).contiguous()
    gmax = int(g.max())
    if gmax >= _MAX_COORD:
        raise ValueError("each batch's coordinate span must be smaller than 65536")
    if max_shift is None:
        max_shift = gmax.bit_length()
    if max_shift < 0 or max_shift > 16:
        raise ValueError("max_shift must lie in [0, 16]")
    h = AdaptiveOctreeHierarchy(grid=g, batch=batch_idx.contiguous(), num_batches=B,
                                coordinate_offsets=mins, max_shift=max_shift)
    pts = None if xyz is None else xyz.float().contiguous()
    for s in range(max_shift + 1):
        gr = Grid().build_cells(g, h.batch, pts, num_clouds=B, num_down=s)
        h.grids.append(gr)
        h.candidate_counts.append(gr.point_candidates(1))
    torch.cuda.synchronize(grid_coord.device)
    h.construction_ms = (time.perf_counter() - t0) * 1e3
    return h


def select_adaptive_levels(hierarchy: AdaptiveOctreeHierarchy, k: int,
                           min_candidates_factor: int = 2,
                           max_candidates_factor: int = 8) -> torch.Tensor:
    """Selected level per point (int16), the reference's coarse-to-fine rule."""
    if k <= 0:
        raise ValueError("k must be positive")
    if min_candidates_factor < 1:
        raise ValueError("min_candidates_factor must be at least one")
    if max_candidates_factor < min_candidates_factor:
        raise ValueError("max_candidates_factor must not be smaller than min")
    n = hierarchy.grid.shape[0]
    if n < k:
        raise ValueError(f"point cloud contains {n} points, fewer than k={k}")
    coarsest = hierarchy.max_shift
    counts = hierarchy.candidate_counts
    selected = torch.full((n,), coarsest, dtype=torch.int16, device=hierarchy.grid.device)
    active = counts[coarsest] > max_candidates_factor * k
    for s in range(coarsest - 1, -1, -1):
        child = counts[s]
        use = active & (child >= min_candidates_factor * k)
        selected[use] = s
        active = use & (child > max_candidates_factor * k)
    return selected


class AdaptiveNeighborhoodFlashKNN:
    """FlashKNN with an independently selected octree level per query point."""

    def __init__(self, num_nbr: int = 32, min_candidates_factor: int = 2,
                 max_candidates_factor: int = 8):
        self.num_nbr = num_nbr
        self.min_candidates_factor = min_candidates_factor
        self.max_candidates_factor = max_candidates_factor
        self.last_stats: dict[str, Any] = {}
        self.last_levels = None      # selected level per point (int32), k <= 64
        self._grids: dict[int, Grid] = {}
        self._ws = None              # workspace of the k <= 64 path

    def _level_grid(self, s: int) -> Grid:
        """Grid of level s, kept between calls (its device tables are reused)."""
        g = self._grids.get(s)
        if g is None:
            g = self._grids[s] = Grid()
        return g

    @torch.no_grad()
    def query(self, grid_coord: torch.Tensor, batch_idx: torch.Tensor,
              xyz: torch.Tensor | None = None, *, max_shift: int | None = None,
              batch_for_prune: int = 1, cut_radius: float = math.inf,
              return_distances: bool = False):
        """k <= 64: two library calls (fk_adaptive_select, fk_adaptive_query)
        over one Morton sort and per-level cell tables, and one host read of
        the statistics at the end.  Larger k: one grid per level (_query_levels)."""
        K = self.num_nbr
        if K < 1 or K > F.FK_MAX_K:
            raise ValueError(f"num_nbr must be in [1, {F.FK_MAX_K}]")
        if K <= F.FK_MAX_BATCHED_K:
            return self._query_device(grid_coord, batch_idx, xyz, max_shift, cut_radius,
                                      return_distances)
        return self._query_levels(grid_coord, batch_idx, xyz, max_shift=max_shift,
                                  cut_radius=cut_radius, return_distances=return_distances)

    def _workspace(self, n, dev):
        nbytes = int(F.load().fk_adaptive_workspace_bytes(n))
        if nbytes < 0:
            raise ValueError(f"adaptive query: {n} points is above the supported size")
        ws = self._ws
        if ws is None or ws.numel() < nbytes or ws.device != dev:
            ws = self._ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        return ws

    def _query_device(self, grid_coord, batch_idx, xyz, max_shift, cut_radius,
                      return_distances):
        K = self.num_nbr
        if not grid_coord.is_cuda or not batch_idx.is_cuda:
            raise ValueError("adaptive octree inputs must be HIP device tensors")
        if grid_coord.dtype != torch.int64 or batch_idx.dtype != torch.int64:
            raise ValueError("grid_coord and batch_idx must use torch.int64")
        if grid_coord.ndim != 2 or grid_coord.shape[1] != 3:
            raise ValueError("grid_coord must have shape [N, 3]")
        if batch_idx.ndim != 1 or batch_idx.shape[0] != grid_coord.shape[0]:
            raise ValueError("batch_idx must have shape [N]")
        n = grid_coord.shape[0]
        if n == 0:
            raise ValueError("cannot construct an octree for an empty point cloud")
        if max_shift is not None and (max_shift < 0 or max_shift > 16):
            raise ValueError("max_shift must lie in [0, 16]")
        if n < K:
            raise ValueError(f"point cloud contains {n} points, fewer than k={K}")
        if self.min_candidates_factor < 1:
            raise ValueError("min_candidates_factor must be at least one")
        if self.max_candidates_factor < self.min_candidates_factor:
            raise ValueError("max_candidates_factor must not be smaller than min")
        dev = grid_coord.device
        L = F.load()
        g = grid_coord.contiguous()
        b = batch_idx.contiguous()
        pts = None if xyz is None else xyz.float().contiguous()
        ws = self._workspace(n, dev)
        stats = torch.empty(F.FK_ADAPTIVE_STATS, dtype=torch.int64, device=dev)
        level = torch.empty(n, dtype=torch.int32, device=dev)
        out_i = torch.empty((n, K), dtype=torch.int64, device=dev)
        out_d = torch.empty((n, K), dtype=torch.float32, device=dev) if return_distances else None
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        s = F.stream_ptr()
        ev[0].record()
        F.check(L.fk_adaptive_select(F.ptr(g), F.ptr(b), F.ptr(pts), n, K,
                                     self.min_candidates_factor, self.max_candidates_factor,
                                     -1 if max_shift is None else int(max_shift), F.ptr(ws),
                                     F.ptr(level), F.ptr(stats), s), "fk_adaptive_select")
        ev[1].record()
        F.check(L.fk_adaptive_query(F.ptr(ws), n, K, float(cut_radius) ** 2, F.ptr(stats),
                                    F.ptr(out_i), F.ptr(out_d), s), "fk_adaptive_query")
        ev[2].record()
        st = stats.cpu().tolist()  # the one host read
        if st[0]:
            if st[1] == 1:
                raise ValueError("batch indices must lie in [0, 32768)")
            if st[1] == 2:
                raise ValueError("each batch's coordinate span must be smaller than 65536")
            F.check(int(st[0]), "fk_adaptive_select (device)")
        smax = int(st[3])
        t_build, t_query = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
        levels = []
        for lv in range(smax + 1):
            m, tot, lo, hi = st[4 + 4 * lv: 8 + 4 * lv]
            if m:
                levels.append({"shift": lv, "query_points": int(m), "candidate_mean": tot / m,
                               "candidate_min": int(lo), "candidate_max": int(hi)})
        self.last_levels = level
        self.last_stats = {
            "construction_ms": t_build, "selection_ms": 0.0, "query_ms": t_query,
            "total_ms": t_build + t_query, "k": K,
            "min_candidates_factor": self.min_candidates_factor,
            "max_candidates_factor": self.max_candidates_factor,
            "octree_levels": smax + 1, "levels_built": smax + 1,
            "selected_levels": levels, "query_kernel_launches": 1,
        }
        if return_distances:
            return out_i, out_d
        return out_i

    @torch.no_grad()
    def _query_levels(self, grid_coord: torch.Tensor, batch_idx: torch.Tensor,
                      xyz: torch.Tensor | None = None, *, max_shift: int | None = None,
                      cut_radius: float = math.inf, return_distances: bool = False):
        """Levels are built from the coarsest down while the selection rule
        runs on the device; the descent stops at the first level where no
        point may go finer (finer levels could not be selected), so the
        finest dense boxes are built only where some point needs them.
        Host reads: one for the sizes, one per level (does any point still
        descend), one for the per-level query counts and one for the
        statistics; no per-point work runs on the host."""
        K = self.num_nbr
        dev = grid_coord.device
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        if not grid_coord.is_cuda or not batch_idx.is_cuda:
            raise ValueError("adaptive octree inputs must be HIP device tensors")
        if grid_coord.dtype != torch.int64 or batch_idx.dtype != torch.int64:
            raise ValueError("grid_coord and batch_idx must use torch.int64")
        if grid_coord.ndim != 2 or grid_coord.shape[1] != 3:
            raise ValueError("grid_coord must have shape [N, 3]")
        if batch_idx.ndim != 1 or batch_idx.shape[0] != grid_coord.shape[0]:
            raise ValueError("batch_idx must have shape [N]")
        n = grid_coord.shape[0]
        if n == 0:
            raise ValueError("cannot construct an octree for an empty point cloud")
        bmin, bmax = torch.aminmax(batch_idx)
        lo = torch.stack([bmin, bmax]).cpu().tolist()
        if lo[0] < 0 or lo[1] >= _MAX_BATCH:
            raise ValueError("batch indices must lie in [0, 32768)")
        B = lo[1] + 1
        mins = batch_min(grid_coord, batch_idx, B)
        g = (grid_coord - mins[batch_idx]).contiguous()
        # per-batch spans (one host read): they size every level's boxes
        spans = batch_max(g, batch_idx, B).cpu().numpy()
        gmax = int(spans.max())
        if gmax >= _MAX_COORD:
            raise ValueError("each batch's coordinate span must be smaller than 65536")
        if max_shift is None:
            max_shift = int(gmax).bit_length()
        if max_shift < 0 or max_shift > 16:
            raise ValueError("max_shift must lie in [0, 16]")
        if n < K:
            raise ValueError(f"point cloud contains {n} points, fewer than k={K}")
        batch = batch_idx.contiguous()
        pts = g.float() if xyz is None else xyz.float().contiguous()
        lo_f, hi_f = self.min_candidates_factor * K, self.max_candidates_factor * K
        sel = torch.full((n,), max_shift, dtype=torch.int16, device=dev)
        counts, built = {}, {}
        active = None

        def box_cells(s):
            # cells of all per-batch boxes at level s: (span >> s) + 1 cells
            # per axis (shifted coordinates start at 0; a batch index without
            # points has span -inf after the minimum, and no box)
            ok = spans[:, 0] >= 0
            return int(np.prod((spans[ok] >> s) + 1, axis=1).sum())

        # The coarse levels (small boxes) form one grid: the points once per
        # level, cloud = level slot * B + batch, each cloud its own box, so a
        # query at level s sees exactly its level's 27 nodes.  One build, one
        # candidate count and (below) one attached query serve them all.
        small = [s for s in range(max_shift, -1, -1) if box_cells(s) <= _SYNC_CELLS]
        Lc = len(small)
        if Lc:
            comb = self._level_grid(-1)
            shifts = torch.tensor(small, dtype=torch.int64, device=dev)
            g_rep = (g[None] >> shifts[:, None, None]).reshape(-1, 3)
            b_rep = (batch[None] + B * torch.arange(Lc, device=dev)[:, None]).reshape(-1)
            comb.reserve(n * Lc, sum(box_cells(s) for s in small) + 64, B * Lc, n)
            comb.build_cells(g_rep, b_rep, pts.repeat(Lc, 1), num_clouds=B * Lc,
                             num_down=0, no_sync=True)
            c_rep = comb.point_candidates(1).view(Lc, n)
            for i, s in enumerate(small):
                counts[s] = c_rep[i]
                built[s] = comb
        reserved = {}  # level -> its grid was reserved (no-sync build and attach)
        first_sep = max_shift  # first level built on its own
        if Lc:
            reserved[-1] = True
            sel, active = walk_levels(c_rep, shifts, lo_f, hi_f)
            first_sep = small[-1] - 1
            # a finer level no point can select changes nothing; read the mask
            # back only before a separate (large) box
            if first_sep >= 0 and not bool(active.any()):
                first_sep = -1
        for s in range(first_sep, -1, -1):
            gr = self._level_grid(s)
            cap = box_cells(s) + 64
            # sparse boxes (many cells per point) take the column layout: a
            # synchronising build, but a table of x-y columns instead of the
            # box's cells to fill and scan (twice: build and attach)
            columns = cap > _COLUMN_CELLS_PER_POINT * n
            no_sync = cap <= _RESERVE_CAP and not columns
            if no_sync:  # sized from the host-side spans: no read-back
                gr.reserve(n, cap, B, n)
            gr.build_cells(g, batch, pts, num_clouds=B, num_down=s, no_sync=no_sync,
                           columns=columns)
            counts[s], built[s] = gr.point_candidates(1), gr
            reserved[s] = no_sync
            c = counts[s]
            if active is None:
                active = c > hi_f
            else:
                use = active & (c >= lo_f)
                sel = torch.where(use, torch.full_like(sel, s), sel)
                active = use & (c > hi_f)
            if s > 0 and not bool(active.any()):
                break
        torch.cuda.synchronize(dev)
        t_build = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        L = max_shift + 1
        sel_l = sel.long()
        # One sort by (selected level, candidate count) orders the points by
        # level -- each level's queries are one contiguous slice of the
        # reordered inputs and outputs (the order inside a level is free: the
        # attach sorts queries by cell, rows go back by id) -- and gives the
        # per-level statistics of the selected count: run bounds by
        # searchsorted, sums by a prefix sum, min/max at the ends of each run
        # (a weighted bincount over a few levels funnels every point into a
        # handful of atomics: 0.2 ms)
        zero = torch.zeros(n, dtype=torch.int32, device=dev)
        c_all = torch.stack([counts.get(s, zero) for s in range(L)])
        c_sel = c_all.gather(0, sel_l[None]).squeeze(0).long()
        keys, perm = torch.sort(sel_l * (1 << 32) + c_sel, stable=True)
        pts_p, g_p, batch_p, sel_p = pts[perm], g[perm], batch[perm], sel_l[perm]
        perm32 = perm.to(torch.int32)
        cs = keys & 0xFFFFFFFF
        bounds = torch.searchsorted(keys, torch.arange(L + 1, device=dev) * (1 << 32))
        starts, ends = bounds[:-1], bounds[1:]
        per_level_d = ends - starts
        csum = torch.cat([cs.new_zeros(1), torch.cumsum(cs, 0)])
        sums = csum[ends] - csum[starts]
        lo_c = cs[starts.clamp(max=n - 1)]
        hi_c = cs[(ends - 1).clamp(min=0)]
        host = torch.stack([per_level_d, sums, lo_c, hi_c]).cpu()
        per_level = [int(v) for v in host[0].tolist()]
        out_i_p = torch.empty((n, K), dtype=torch.int64, device=dev)
        out_d_p = torch.empty((n, K), dtype=torch.float32, device=dev)
        qkw = dict(mode=F.FK_MODE_CELLS, radius_cells=1, cut_r2=float(cut_radius) ** 2,
                   pad_mode=F.FK_PAD_PER_QUERY, attached=True)
        start, used, stats = 0, [], []
        s_small = small[-1] if Lc else L  # levels >= s_small are in the combined grid
        for s in range(L):
            m = per_level[s]
            if m:
                used.append((s, m))
                stats.append((float(host[1][s]) / m, int(host[2][s]), int(host[3][s])))
            if m == 0 or s >= s_small:
                continue
            sl = slice(start, start + m)
            start += m
            gr = built[s]
            gr.attach_queries(pts_p[sl], qcell=g_p[sl], qbatch=batch_p[sl],
                              qpad=perm32[sl], no_sync=reserved[s])
            gr.query(K, out_idx=out_i_p[sl], out_dist=out_d_p[sl], **qkw)
        if start < n:  # the combined coarse levels: one attach, one launch
            sl = slice(start, n)
            sq = sel_p[sl]
            comb = built[s_small]
            comb.attach_queries(pts_p[sl], qcell=g_p[sl] >> sq[:, None],
                                qbatch=batch_p[sl] + B * (max_shift - sq),
                                qpad=perm32[sl], no_sync=True)
            comb.query(K, out_idx=out_i_p[sl], out_dist=out_d_p[sl], **qkw)
            # support ids of level slot i are offset by i * n; pads are query ids < n
            out_i_p[sl] = out_i_p[sl] % n
        out_i = torch.empty_like(out_i_p)
        out_d = torch.empty_like(out_d_p)
        out_i[perm] = out_i_p
        out_d[perm] = out_d_p
        for gr in {id(v): v for v in built.values()}.values():  # device errors of
            gr.status()                                         # the no-sync work
        t_query = (time.perf_counter() - t0) * 1e3
        self.last_stats = {
            "construction_ms": t_build, "selection_ms": 0.0,
            "query_ms": t_query, "total_ms": t_build + t_query,
            "k": K, "min_candidates_factor": self.min_candidates_factor,
            "max_candidates_factor": self.max_candidates_factor,
            "octree_levels": max_shift + 1, "levels_built": len(counts),
            "selected_levels": [
                {"shift": int(s), "query_points": int(m), "candidate_mean": st[0],
                 "candidate_min": int(st[1]), "candidate_max": int(st[2])}
                for (s, m), st in zip(used, stats)],
            "query_kernel_launches": sum(1 for s, _ in used if s < s_small) + int(start < n),
        }
        if return_distances:
            return out_i, out_d
        return out_i
