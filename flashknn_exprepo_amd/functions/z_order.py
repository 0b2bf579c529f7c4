"""Morton keys (drop-in for FlashKNN/functions/z_order.py:57-91).

``xyz2key(x, y, z, b=None, depth=16)`` returns the shuffled key with bit i
of x/y/z at bit 3i+2/3i+1/3i of the key and ``b << 48`` on top; coordinates
keep their low ``depth`` bits (the reference's KeyLUT masks the same way,
z_order.py:81-86).  On a HIP device with depth 16 the key is computed by
the library kernel (the fused path of FlashKnnWrapper.py:220-229); other
cases use the same bit layout with torch integer ops.
"""
from __future__ import annotations

from typing import Optional, Union

import torch

from .. import _lib as F


def _spread3(v: torch.Tensor) -> torch.Tensor:
    # standard magic-number interleave: bit i -> bit 3i (16 input bits)
    v = v & 0xFFFF
    v = (v | (v << 16)) & 0x0000FF0000FF
    v = (v | (v << 8)) & 0x00F00F00F00F
    v = (v | (v << 4)) & 0x0C30C30C30C3
    v = (v | (v << 2)) & 0x249249249249
    return v


def xyz2key(
    x: torch.Tensor,
    y: torch.Tensor,
    z: torch.Tensor,
    b: Optional[Union[torch.Tensor, int]] = None,
    depth: int = 16,
) -> torch.Tensor:
    if depth < 1 or depth > 16:
        raise ValueError("depth must be in [1, 16]")
    x, y, z = x.long(), y.long(), z.long()
    if depth == 16 and x.is_cuda and (b is None or isinstance(b, torch.Tensor)):
        grid = torch.stack([x, y, z], dim=-1).reshape(-1, 3).contiguous()
        bb = None
        if b is not None:
            bb = b.long().expand_as(x).reshape(-1).contiguous()
        keys = torch.empty(grid.shape[0], dtype=torch.int64, device=x.device)
        L = F.load()
        F.check(L.fk_morton_encode(F.ptr(grid), F.ptr(bb), grid.shape[0], F.ptr(keys),
                                   F.stream_ptr()), "fk_morton_encode")
        return keys.reshape(x.shape)
    mask = (1 << depth) - 1
    key = (_spread3(x & mask) << 2) | (_spread3(y & mask) << 1) | _spread3(z & mask)
    if b is not None:
        key = (torch.as_tensor(b, device=key.device).long() << 48) | key
    return key


def neighbor_keys(grid_down: torch.Tensor, parent_batch: Optional[torch.Tensor] = None
                  ) -> torch.Tensor:
    """27 neighbour keys per parent, offsets [0,-1,1]^3 (centre first); drop-in
    for FlashKNN_Morton_Neighbor_Keys (flash_knn_construction.cu:105-126)."""
    g = grid_down.long().contiguous()
    F.require_device(g, "grid_down")
    pb = None if parent_batch is None else parent_batch.long().contiguous()
    out = torch.empty(g.shape[0] * 27, dtype=torch.int64, device=g.device)
    L = F.load()
    F.check(L.fk_morton_neighbor_keys(F.ptr(g), F.ptr(pb), g.shape[0], F.ptr(out),
                                      F.stream_ptr()), "fk_morton_neighbor_keys")
    return out
