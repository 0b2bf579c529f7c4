"""Voxel-first subsampling and the FlashKNN level hierarchy on the device.

``voxel_first(xyz, size, offsets=None)`` replaces _voxel_first_cuda
(Networks/hierarchy.py:23-36): for every occupied voxel floor(xyz / size)
(correctly rounded fp32 quotient; voxel coordinates keyed by their low 16
bits per axis as xyz2key does) it keeps the point with the smallest index
and returns the kept indices ascending.  One C-ABI call (fk_voxel_first: hash insert with
atomicMin, flag scan, compaction); the only host read is the kept count.

``build_flash_hierarchy(metric_xyz, voxel_sizes, ks, alpha)`` mirrors
Networks/hierarchy.py:39-97: per level, voxel-first subsampling of the
previous level, a FlashKNN.query on cells of ``size * alpha``, then the
back queries mapping every level-0 point to each coarser level, returned as
``forward + backward`` in the reference's order.
"""
from __future__ import annotations

import math

import torch

from . import _lib as F
from .functions.FlashKnnWrapper import FlashKNN, quotient32


def voxel_first(xyz: torch.Tensor, size: float, offsets: torch.Tensor = None) -> torch.Tensor:
    """int64 indices (ascending) of each voxel's first point; ``offsets``
    (cumulative int32 ends) voxelises every cloud on its own."""
    F.require_device(xyz, "xyz")
    L = F.load()
    xyz = xyz.float().contiguous()
    n = xyz.shape[0]
    off = None
    B = 1
    if offsets is not None:
        off = offsets.to(device=xyz.device, dtype=torch.int32).contiguous()
        B = off.shape[0]
    nbytes = L.fk_voxel_workspace_bytes(n)
    if nbytes < 0:
        raise ValueError(f"voxel_first: {n} points is above the supported size")
    ws = torch.empty(nbytes, dtype=torch.uint8, device=xyz.device)
    out = torch.empty(n, dtype=torch.int64, device=xyz.device)
    cnt = torch.empty(1, dtype=torch.int64, device=xyz.device)
    F.check(L.fk_voxel_first(F.ptr(xyz) if n else None, F.ptr(off), B, n, float(size),
                             F.ptr(ws), F.ptr(out) if n else None, F.ptr(cnt),
                             F.stream_ptr()), "fk_voxel_first")
    c = int(cnt.item())
    if c < 0:
        F.check(-c, "fk_voxel_first (device)")
    return out[:c]


def build_flash_hierarchy(metric_xyz: torch.Tensor,
                          voxel_sizes=(0.06, 0.12, 0.24, 0.48),
                          ks=(24, 24, 24, 24), alpha: int = 16):
    """Neighbour and back-mapping indices of the DeLA level hierarchy
    (Networks/hierarchy.py:39-97), in the order the reference returns them."""
    exponent = round(math.log2(alpha)) if alpha > 0 else 0
    if alpha < 2 or 2 ** exponent != alpha:
        raise ValueError("alpha must be a power of two")
    if len(voxel_sizes) != len(ks):
        raise ValueError("voxel_sizes and ks must have equal length")
    F.require_device(metric_xyz, "metric_xyz")
    full_xyz = metric_xyz
    current = metric_xyz
    dev = metric_xyz.device
    cumulative = torch.arange(len(metric_xyz), device=dev)
    batch = torch.zeros(len(metric_xyz), device=dev, dtype=torch.long)
    forward, levels = [], []
    for level, (size, k) in enumerate(zip(voxel_sizes, ks)):
        if level:
            local = voxel_first(current, size)
            current = current[local].contiguous()
            cumulative = cumulative[local]
            batch = batch[local]
            forward.append(local)
            levels.append(cumulative)
        grid = torch.floor(quotient32(current, size)).long().contiguous()
        forward.append(FlashKNN(num_nbr=k, num_down=exponent).query(
            grid, batch, current, memory_mode="SM", sorting_mode="PS"))
    backward = []
    full_batch = torch.zeros(len(full_xyz), device=dev, dtype=torch.long)
    for level in range(len(levels) - 1, -1, -1):
        size = voxel_sizes[level + 1]
        backward.append(FlashKNN(num_nbr=1).back_query(
            full_xyz, levels[level], query_grid_size=size * 2, down_grid_size=size,
            batch_idx=full_batch))
    return forward + backward
