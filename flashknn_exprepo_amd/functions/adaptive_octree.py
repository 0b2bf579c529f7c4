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

import torch

from .. import _lib as F
from ..grid import Grid

_MAX_COORD = 1 << 16
_MAX_BATCH = 32768


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
    mins = torch.full((B, 3), torch.iinfo(torch.int64).max, dtype=torch.int64,
                      device=grid_coord.device)
    mins.scatter_reduce_(0, batch_idx[:, None].expand(-1, 3), grid_coord, "amin",
                         include_self=True)
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

    @torch.no_grad()
    def query(self, grid_coord: torch.Tensor, batch_idx: torch.Tensor,
              xyz: torch.Tensor | None = None, *, max_shift: int | None = None,
              batch_for_prune: int = 1, cut_radius: float = math.inf,
              return_distances: bool = False):
        K = self.num_nbr
        if K < 1 or K > F.FK_MAX_K:
            raise ValueError(f"num_nbr must be in [1, {F.FK_MAX_K}]")
        h = build_adaptive_octree(grid_coord, batch_idx, max_shift, xyz)
        t0 = time.perf_counter()
        sel = select_adaptive_levels(h, K, self.min_candidates_factor,
                                     self.max_candidates_factor)
        torch.cuda.synchronize(grid_coord.device)
        t_sel = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        n = grid_coord.shape[0]
        dev = grid_coord.device
        out_i = torch.empty((n, K), dtype=torch.int64, device=dev)
        out_d = torch.empty((n, K), dtype=torch.float32, device=dev)
        pts = h.grid.float() if xyz is None else xyz.float()
        levels_used = torch.unique(sel).tolist()
        level_stats = []
        for s in levels_used:
            q = torch.nonzero(sel == s).flatten()
            gr = h.grids[s]
            gr.attach_queries(pts[q].contiguous(), qcell=h.grid[q], qbatch=h.batch[q],
                              qpad=q.to(torch.int32))
            qi, qd = gr.query(K, mode=F.FK_MODE_CELLS, radius_cells=1,
                              cut_r2=float(cut_radius) ** 2,
                              pad_mode=F.FK_PAD_PER_QUERY, attached=True)
            out_i[q] = qi
            out_d[q] = qd
            cc = h.candidate_counts[s][q]
            level_stats.append({"shift": int(s), "query_points": int(q.numel()),
                                "candidate_mean": float(cc.float().mean()),
                                "candidate_min": int(cc.min()),
                                "candidate_max": int(cc.max())})
        torch.cuda.synchronize(dev)
        t_query = (time.perf_counter() - t0) * 1e3
        self.last_stats = {
            "construction_ms": h.construction_ms, "selection_ms": t_sel,
            "query_ms": t_query, "total_ms": h.construction_ms + t_sel + t_query,
            "k": K, "min_candidates_factor": self.min_candidates_factor,
            "max_candidates_factor": self.max_candidates_factor,
            "octree_levels": h.max_shift + 1, "selected_levels": level_stats,
            "query_kernel_launches": len(levels_used),
        }
        if return_distances:
            return out_i, out_d
        return out_i
