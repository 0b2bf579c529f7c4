"""Exact kNN on the grid (certified), the engine behind the pointops facade.

``knn(xyz, offsets, k, new_xyz=None, new_offsets=None)`` returns the exact
k nearest support points of every query within its cloud (fp32 squared
distance ((dx*dx + dy*dy) + dz*dz), ties to the smaller index), the result
pointops.knn_query computes by scanning whole segments
(Pointcept/libs/pointops/src/knn_query/knn_query_cuda_kernel.cu:60-104).

Cell size: a first guess from each cloud's bounding-box density
(fk_grid_auto_cell), then density-feedback rebuilds (fk_grid_adapt_cell)
that move the mean 27-cell candidate count towards ``target_ratio * k``.
Exactness never depends on the cell size, only speed does; the target
ratio comes from the autotuner table, keyed by cloud size, k and a density
class measured at the first guess.  Builds use reserved capacity and no
host synchronisation; one status read at the end of a call reports errors
(and a cell box that outgrew the reserve is rebuilt with exact sizing).
"""
from __future__ import annotations

import math

import torch

from . import _lib as F
from .grid import Grid

# first guess: cell side = scale * (k * bbox_volume / n)^(1/3)
DEFAULT_CELL_SCALE = 1.0
# feedback: aim at target_ratio * k candidates per query in the 27 cells
# (9: measured best for the 100k room at k = 16 and 32 on MI355X)
DEFAULT_TARGET_RATIO = 9.0
CANDIDATE_EXPONENT = 2.0  # candidates ~ side^2 on scanned surfaces
ADAPT_ROUNDS = 2


def cloud_ids(offsets: torch.Tensor, n: int) -> torch.Tensor:
    """Cloud id of every point from cumulative end offsets (no host sync)."""
    ar = torch.arange(n, device=offsets.device, dtype=offsets.dtype)
    return torch.searchsorted(offsets.contiguous(), ar, right=True)


def _cell_sizes_torch(xyz, offsets, k, scale):
    """Host restatement of fk_grid_auto_cell for CPU tensors."""
    B = offsets.shape[0]
    cid = cloud_ids(offsets, xyz.shape[0]).long()
    big = torch.full((B, 3), float("inf"), device=xyz.device)
    mn = big.scatter_reduce(0, cid[:, None].expand(-1, 3), xyz, "amin", include_self=True)
    mx = (-big).scatter_reduce(0, cid[:, None].expand(-1, 3), xyz, "amax", include_self=True)
    ends = offsets.long()
    starts = torch.cat([ends.new_zeros(1), ends[:-1]])
    n = (ends - starts).clamp_min(1).double()
    ext = (mx.double() - mn.double()).clamp_min(0)
    ext = torch.where(torch.isfinite(ext), ext, torch.zeros_like(ext))
    span = ext.max(1).values.clamp_min(1e-6)
    vol = ext.clamp_min(1e-2 * span[:, None]).prod(1)
    kk = torch.as_tensor(k, device=xyz.device).double().expand(B)
    side = float(scale) * (kk * vol / n).pow(1.0 / 3.0)
    side = torch.maximum(side, (vol / (64.0 * n)).pow(1.0 / 3.0)).clamp_min(1e-6)
    return side.float()


_util = {}


def cell_sizes(xyz: torch.Tensor, offsets: torch.Tensor, k, scale=None) -> torch.Tensor:
    """Per-cloud first-guess cell side from the bounding-box density.

    On a HIP device this runs fk_grid_auto_cell (no host sync); CPU tensors
    use the identical torch formula."""
    if scale is None:
        scale = DEFAULT_CELL_SCALE
    if not xyz.is_cuda:
        return _cell_sizes_torch(xyz.float(), offsets, k, scale)
    dev = torch.cuda.current_device()
    g = _util.get(dev)
    if g is None:
        g = _util[dev] = Grid()
    side, _ = g.auto_cell(xyz, offsets, k, scale)
    return side


ADAPT_MIN_POINTS = 16384   # smaller clouds keep the first guess
# One-shot calls (build, one query) skip the feedback rebuilds when every
# cloud's first guess is within this factor of its candidate target: there
# the two rebuilds cost more than the query saves (MI355X, build + one K=16
# query: room 250k 563 -> 460 us, uniform 1M 1392 -> 896 us, the hierarchy's
# 28k -> 98k stage 445 -> 332 us), while uneven clouds far off target keep
# them (LiDAR 92k, first guess 13x the target: 546 us, 784 without).
ONE_SHOT_BAND = 4.0


def density_class(mean_candidates, k) -> int:
    """log2 of (first-guess 27-cell candidates per query) / k: how far the
    points concentrate below their bounding-box density (surfaces, LiDAR
    rings); part of the autotuner's bucket key."""
    r = max(float(mean_candidates) / max(int(k), 1), 1.0)
    return max(0, min(15, int(math.floor(math.log2(r)))))


def build_exact_grid(g: Grid, xyz, offsets, k, *, cell=None, scale=None,
                     target_ratio=None, adapt=True, rounds=None, no_sync=True,
                     prune=None, skip_band=None, min_points=ADAPT_MIN_POINTS):
    """Build ``g`` for exact queries of (xyz, offsets); returns the per-cloud
    cell side.  First guess from the bounding-box density, then (clouds of
    >= ADAPT_MIN_POINTS points) density-feedback rebuilds towards the
    autotuned ``target_ratio * k`` candidates per query.  With ``no_sync``
    the builds use the reserved capacity and read nothing back except the
    per-cloud first-guess statistic (one read).  With ``skip_band`` b the
    rebuilds are skipped when every cloud's first-guess candidates per query
    lie within a factor b of its target.  Sets ``g.prune_hint``
    (FK_QUERY_PRUNE for the queries on this build): ``prune`` if given, else
    the autotuner's choice for the clouds holding most points; likewise
    ``g.one_lane_hint`` and ``g.batch_hint`` (the candidate batch).  Builds of
    fewer than ``min_points`` points keep the first guess."""
    n = xyz.shape[0]
    g.prune_hint = bool(prune) if prune is not None else False
    g.one_lane_hint = False
    g.batch_hint = "default"
    if cell is not None:
        side = torch.as_tensor(cell, dtype=torch.float32, device=xyz.device).reshape(-1)
        side = side.expand(offsets.shape[0]).contiguous()
        g.build_xyz(xyz, offsets, (1.0 / side.double()).float(), no_sync=no_sync)
        return side
    side, inv = g.auto_cell(xyz, offsets, k, DEFAULT_CELL_SCALE if scale is None else scale)
    g.build_xyz(xyz, offsets, inv, no_sync=no_sync)
    if not adapt or n < min_points:
        return side
    if target_ratio is None:
        from .autotune import (lookup_batch, lookup_one_lane, lookup_prune,
                               lookup_target_ratio)
        ends = offsets.long()
        sizes = torch.diff(ends, prepend=ends.new_zeros(1)).clamp_min(1)
        # one host read: per-cloud sizes and first-guess candidates per query
        both = torch.stack([sizes.double(), g.cloud_candidates().double() / sizes.double()])
        ns, mean = both.cpu().tolist()
        ns = [int(v) for v in ns]
        ks = k.tolist() if isinstance(k, torch.Tensor) else [int(k)] * len(mean)
        dcs = [density_class(mc, kc) for kc, mc in zip(ks, mean)]
        ratios = [lookup_target_ratio(nc, kc, dc) for nc, kc, dc in zip(ns, ks, dcs)]
        target_ratio = torch.tensor(ratios, dtype=torch.float32, device=xyz.device)
        def vote(look):  # the choice of the clouds holding most points
            return sum(nc if look(nc, kc, dc) else -nc
                       for nc, kc, dc in zip(ns, ks, dcs)) > 0
        if prune is None:
            g.prune_hint = vote(lookup_prune)
        g.one_lane_hint = vote(lookup_one_lane)
        weight = {}
        for nc, kc, dc in zip(ns, ks, dcs):
            b = lookup_batch(nc, kc, dc)
            weight[b] = weight.get(b, 0) + nc
        g.batch_hint = max(weight, key=lambda b: (weight[b], b == "default"))
        if skip_band is not None and all(
                1.0 / skip_band <= mc / (r * kc) <= skip_band
                for mc, r, kc in zip(mean, ratios, ks)):
            return side
    for _ in range(ADAPT_ROUNDS if rounds is None else rounds):
        g.adapt_cell(k, side, inv, target_ratio, CANDIDATE_EXPONENT)
        g.build_xyz(xyz, offsets, inv, no_sync=no_sync)
    return side


class ExactKNN:
    """Reusable exact-kNN engine (keeps its device tables between calls)."""

    def __init__(self):
        self.grid = Grid()
        self.last_stats = None
        self.side = None
        self._cap = (0, 0, 0)

    def _reserve(self, n, B, nq):
        need = (n, B, nq)
        if all(a <= c for a, c in zip(need, self._cap)):
            return
        n2, B2, q2 = (max(int(a * 1.25) + 16, c) for a, c in zip(need, self._cap))
        cells = min(128 * n2 + 16 * B2 + 65536, 1 << 28)
        self.grid.reserve(n2, cells, B2, max(n2, q2))
        self._cap = (n2, B2, q2)

    def build(self, xyz, offsets, k, *, cell=None, scale=None, target_ratio=None,
              adapt=True, nq=0, no_sync=None, prune=None, skip_band=None,
              min_points=ADAPT_MIN_POINTS):
        """Build the support grid; returns the per-cloud cell side.  The
        FK_QUERY_PRUNE choice for queries on it is ``self.grid.prune_hint``.

        ``no_sync`` (default: only with automatic cell sizes) builds into
        reserved capacity without a host read.  Automatic sizes keep every
        box below the reserve (at most ~64 cells per point); a caller's
        ``cell`` may not, so by default such a build sizes its box exactly.
        A no-sync build that overflows is reported by the next synchronising
        call (``grid.status()``), and queries on it write nothing.
        ``skip_band``: see ``build_exact_grid`` (one-shot calls pass
        ONE_SHOT_BAND; a grid that serves many queries keeps the rebuilds)."""
        F.require_device(xyz, "xyz")
        if no_sync is None:
            no_sync = cell is None
        xyz = xyz.float().contiguous()
        offsets = offsets.to(device=xyz.device, dtype=torch.int32).contiguous()
        if no_sync:
            self._reserve(xyz.shape[0], offsets.shape[0], max(nq, xyz.shape[0]))
        self.side = build_exact_grid(self.grid, xyz, offsets, k, cell=cell, scale=scale,
                                     target_ratio=target_ratio, adapt=adapt,
                                     no_sync=no_sync, prune=prune, skip_band=skip_band,
                                     min_points=min_points)
        return self.side

    @torch.no_grad()
    def __call__(self, xyz, offsets, k, new_xyz=None, new_offsets=None, *,
                 idx_dtype=torch.int64, pad_mode=F.FK_PAD_MINUS1,
                 dist_mode=F.FK_DIST_SQ, pad_dist=math.inf, cell=None, scale=None,
                 target_ratio=None, adapt=True, cut_r2=math.inf, stats=False):
        if k < 1 or k > F.FK_MAX_K:
            raise ValueError(f"k must be in [1, {F.FK_MAX_K}], got {k}")
        attached = new_xyz is not None
        if attached:
            if new_offsets is None:
                raise ValueError("new_offsets required with new_xyz")
            F.require_device(new_xyz, "new_xyz")
        if pad_mode == F.FK_PAD_SELF and attached:
            pad_mode = F.FK_PAD_MINUS1
        nq = new_xyz.shape[0] if attached else 0
        try:
            out = self._run(xyz, offsets, k, new_xyz, new_offsets, attached, nq,
                            dict(cell=cell, scale=scale, target_ratio=target_ratio,
                                 adapt=adapt, skip_band=ONE_SHOT_BAND),
                            dict(cut_r2=cut_r2, idx_dtype=idx_dtype, pad_mode=pad_mode,
                                 dist_mode=dist_mode, pad_dist=pad_dist), True)
        except F.FlashKNNError as e:
            if e.status != F.FK_ERR_GRID_TOO_LARGE:
                raise
            # the dense boxes outgrew the reserved capacity: size them exactly
            out = self._run(xyz, offsets, k, new_xyz, new_offsets, attached, nq,
                            dict(cell=cell, scale=scale, target_ratio=target_ratio,
                                 adapt=adapt, skip_band=ONE_SHOT_BAND),
                            dict(cut_r2=cut_r2, idx_dtype=idx_dtype, pad_mode=pad_mode,
                                 dist_mode=dist_mode, pad_dist=pad_dist), False)
        if stats:
            self.last_stats = self.grid.query_stats()
        return out

    def _run(self, xyz, offsets, k, new_xyz, new_offsets, attached, nq, bkw, qkw, no_sync):
        self.build(xyz, offsets, k, nq=nq, no_sync=no_sync, **bkw)
        g = self.grid
        if attached:
            g.attach_queries(new_xyz.float().contiguous(),
                             qoffsets=new_offsets.to(device=xyz.device, dtype=torch.int32),
                             no_sync=no_sync)
        out = g.query(k, mode=F.FK_MODE_EXACT, radius_cells=1, attached=attached,
                      **g.hints(), **qkw)
        g.status()  # the one host read: reports any device-side error
        return out


class StaticKNN:
    """Exact kNN for a step that repeats on inputs of fixed shape (a
    training step's level-to-level queries), capturable into a graph.

    The first call runs ``ExactKNN`` eagerly (density feedback, host reads)
    and freezes the per-cloud cell sizes and speed hints it chose.  Later
    calls build with those sizes into the reserved capacity and query
    attached points with the counter reset included (no host read, no new
    device table once the first static call has sized them), so a
    ``torch.cuda.graph`` can capture them; outputs are the same exact kNN.
    Errors of captured calls (a box that outgrew the reserve) surface at
    ``status()``."""

    def __init__(self):
        self.eng = ExactKNN()
        self.inv = None

    def status(self):
        self.eng.grid.status()

    @torch.no_grad()
    def __call__(self, xyz, offsets, k, new_xyz=None, new_offsets=None, *,
                 idx_dtype=torch.int64, pad_mode=F.FK_PAD_MINUS1):
        if self.inv is None:
            out = self.eng(xyz, offsets, k, new_xyz, new_offsets, idx_dtype=idx_dtype,
                           pad_mode=pad_mode)
            self.inv = (1.0 / self.eng.side.double()).float()
            self.hints = self.eng.grid.hints()
            return out
        F.require_device(xyz, "xyz")
        xyz = xyz.float().contiguous()
        offsets = offsets.to(device=xyz.device, dtype=torch.int32).contiguous()
        attached = new_xyz is not None
        nq = new_xyz.shape[0] if attached else 0
        self.eng._reserve(xyz.shape[0], offsets.shape[0], max(nq, xyz.shape[0]))
        g = self.eng.grid
        g.build_xyz(xyz, offsets, self.inv, no_sync=True)
        if attached:
            g.attach_queries(new_xyz.float().contiguous(),
                             qoffsets=new_offsets.to(device=xyz.device, dtype=torch.int32),
                             no_sync=True)
        return g.query(k, mode=F.FK_MODE_EXACT, radius_cells=1, attached=attached,
                       idx_dtype=idx_dtype, pad_mode=pad_mode, passes=7, **self.hints)


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
