"""Exact kNN on the grid (certified), the engine behind the pointops facade.

``knn(xyz, offsets, k, new_xyz=None, new_offsets=None)`` returns the exact
k nearest support points of every query within its cloud (fp32 squared
distance ((dx*dx + dy*dy) + dz*dz), ties to the smaller index), the result
pointops.knn_query computes by scanning whole segments
(Pointcept/libs/pointops/src/knn_query/knn_query_cuda_kernel.cu:60-104).
The grid's cell size per cloud comes from ``cell_sizes`` (autotuner table or
density heuristic); exactness never depends on it, only speed does.
"""
from __future__ import annotations

import math

import torch

from . import _lib as F
from .grid import Grid

# cell side = scale * (k * bbox_volume / n)^(1/3); tuned per (n, k, shape)
DEFAULT_CELL_SCALE = 0.6


def cloud_ids(offsets: torch.Tensor, n: int) -> torch.Tensor:
    """Cloud id of every point from cumulative end offsets (no host sync)."""
    ar = torch.arange(n, device=offsets.device, dtype=offsets.dtype)
    return torch.searchsorted(offsets.contiguous(), ar, right=True)


def cell_sizes(xyz: torch.Tensor, offsets: torch.Tensor, k, scale=None) -> torch.Tensor:
    """Per-cloud cell side from the bounding-box density, on the device.

    ``k`` is an int or a per-cloud tensor; ``scale`` an override of the
    autotuned factor.  Returns float32 [num_clouds] (no host sync)."""
    if scale is None:
        from .autotune import lookup_scale
        scale = lookup_scale(xyz.shape[0], k if isinstance(k, int) else int(k.max()))
    B = offsets.shape[0]
    cid = cloud_ids(offsets, xyz.shape[0]).long()
    big = torch.full((B, 3), float("inf"), device=xyz.device)
    mn = big.scatter_reduce(0, cid[:, None].expand(-1, 3), xyz, "amin", include_self=True)
    mx = (-big).scatter_reduce(0, cid[:, None].expand(-1, 3), xyz, "amax", include_self=True)
    ends = offsets.long()
    starts = torch.cat([ends.new_zeros(1), ends[:-1]])
    n = (ends - starts).clamp_min(1).double()
    ext = (mx - mn).double().clamp_min(0)
    ext = torch.where(torch.isfinite(ext), ext, torch.zeros_like(ext))
    span = ext.max(1).values.clamp_min(1e-6)
    # degenerate axes (planar / linear clouds) count with 1% of the span
    vol = ext.clamp_min(1e-2 * span[:, None]).prod(1)
    kk = torch.as_tensor(k, device=xyz.device).double().expand(B)
    side = float(scale) * (kk * vol / n).pow(1.0 / 3.0)
    # keep dense boxes bounded: at most ~64 cells per point per cloud
    min_side = (vol / (64.0 * n)).pow(1.0 / 3.0)
    side = torch.maximum(side, min_side).clamp_min(1e-6)
    return side.float()


class ExactKNN:
    """Reusable exact-kNN engine (keeps its device tables between calls)."""

    def __init__(self):
        self.grid = Grid()
        self.last_stats = None

    @torch.no_grad()
    def __call__(self, xyz, offsets, k, new_xyz=None, new_offsets=None, *,
                 idx_dtype=torch.int64, pad_mode=F.FK_PAD_MINUS1,
                 dist_mode=F.FK_DIST_SQ, pad_dist=math.inf, cell=None, scale=None,
                 cut_r2=math.inf, stats=False):
        if k < 1 or k > F.FK_MAX_K:
            raise ValueError(f"k must be in [1, {F.FK_MAX_K}], got {k}")
        F.require_device(xyz, "xyz")
        xyz = xyz.float().contiguous()
        offsets = offsets.to(device=xyz.device, dtype=torch.int32).contiguous()
        if cell is None:
            cell = cell_sizes(xyz, offsets, k, scale)
        inv = (1.0 / cell.double()).float().to(xyz.device)
        g = self.grid
        g.build_xyz(xyz, offsets, inv)
        attached = new_xyz is not None
        if attached:
            if new_offsets is None:
                raise ValueError("new_offsets required with new_xyz")
            g.attach_queries(new_xyz.float().contiguous(),
                             qoffsets=new_offsets.to(device=xyz.device, dtype=torch.int32))
        if pad_mode == F.FK_PAD_SELF and attached:
            pad_mode = F.FK_PAD_MINUS1
        out = g.query(k, mode=F.FK_MODE_EXACT, radius_cells=1, cut_r2=cut_r2,
                      idx_dtype=idx_dtype, pad_mode=pad_mode, dist_mode=dist_mode,
                      pad_dist=pad_dist, attached=attached)
        if stats:
            self.last_stats = g.query_stats()
        return out


_engines = {}


def _engine() -> ExactKNN:
    dev = torch.cuda.current_device()
    e = _engines.get(dev)
    if e is None:
        e = _engines[dev] = ExactKNN()
    return e


def knn(xyz, offsets, k, new_xyz=None, new_offsets=None, **kw):
    """Exact kNN: (idx [M, k] int64, squared distance [M, k] float32)."""
    F.require_device(xyz, "xyz")
    with torch.cuda.device(xyz.device):
        return _engine()(xyz, offsets, k, new_xyz, new_offsets, **kw)
