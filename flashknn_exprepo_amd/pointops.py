"""pointops facade (Pointcept/libs/pointops/functions/query.py,
interpolation.py, utils.py) on the exact grid kNN.

* ``knn_query(nsample, xyz, offset, new_xyz=None, new_offset=None)`` ->
  (idx int32 [m, nsample], dist float32 [m, nsample]) with the reference's
  conventions (query.py:7-24, knn_query_cuda_kernel.cu:60-104): offsets are
  cumulative segment ends, idx are global indices into xyz, dist is the
  square root of the squared distance, missing slots are (-1, sqrt(1e10)).
  Results are exact kNN with ties to the smaller index (the reference's heap
  order among equal distances is not specified).  nsample <= 128, the
  reference's limit (knn_query_cuda_kernel.cu:82-83); above 64 each query
  runs on a whole wave.
* ``interpolation`` / ``interpolation2`` (interpolation.py:8-59): inverse-
  distance weights 1/(dist + 1e-8) over the k nearest, normalised; the
  backward returns the feature gradient through a deterministic segmented
  sum (the reference scatters with atomicAdd).
* ``offset2batch`` / ``batch2offset`` (utils.py:102-121).
"""
from __future__ import annotations

import torch
from torch.autograd import Function

from . import _lib as F
from .knn import knn


class KNNQuery(Function):
    @staticmethod
    def forward(ctx, nsample, xyz, offset, new_xyz=None, new_offset=None):
        if new_xyz is None or new_offset is None:
            new_xyz, new_offset = xyz, offset
        assert xyz.is_contiguous() and new_xyz.is_contiguous()
        if new_xyz is xyz and new_offset is offset:
            idx, dist = knn(xyz, offset, nsample, idx_dtype=torch.int32,
                            pad_mode=F.FK_PAD_MINUS1, dist_mode=F.FK_DIST_SQRT,
                            pad_dist=1e5)
        else:
            idx, dist = knn(xyz, offset, nsample, new_xyz, new_offset,
                            idx_dtype=torch.int32, pad_mode=F.FK_PAD_MINUS1,
                            dist_mode=F.FK_DIST_SQRT, pad_dist=1e5)
        ctx.mark_non_differentiable(idx, dist)
        return idx, dist

    @staticmethod
    def backward(ctx, a=None, b=None):
        return None, None, None, None, None


knn_query = KNNQuery.apply


def _idw_weights(dist):
    r = 1.0 / (dist + 1e-8)
    return r / r.sum(dim=1, keepdim=True)


def interpolation(xyz, new_xyz, feat, offset, new_offset, k=3):
    """Weighted gather with torch ops (interpolation.py:8-22)."""
    assert xyz.is_contiguous() and new_xyz.is_contiguous() and feat.is_contiguous()
    idx, dist = knn_query(k, xyz, offset, new_xyz, new_offset)
    w = _idw_weights(dist)
    out = torch.zeros(new_xyz.shape[0], feat.shape[1], device=feat.device, dtype=feat.dtype)
    for i in range(k):
        out += feat[idx[:, i].long(), :] * w[:, i].unsqueeze(-1)
    return out


class Interpolation(Function):
    @staticmethod
    def forward(ctx, xyz, new_xyz, input, offset, new_offset, k=3):
        assert xyz.is_contiguous() and new_xyz.is_contiguous() and input.is_contiguous()
        idx, dist = knn_query(k, xyz, offset, new_xyz, new_offset)
        w = _idw_weights(dist).contiguous()
        idx64 = idx.long().contiguous()
        from .autograd import idw_gather_forward
        out = idw_gather_forward(input.float().contiguous(), idx64, w)
        ctx.m, ctx.k = input.shape[0], k
        ctx.save_for_backward(idx64, w)
        return out

    @staticmethod
    def backward(ctx, grad_output):
        idx, w = ctx.saved_tensors
        from .autograd import idw_gather_backward
        grad_input, _ = idw_gather_backward(None, ctx.m, idx, w,
                                            grad_output.float().contiguous(),
                                            need_w=False)
        return None, None, grad_input, None, None, None


interpolation2 = Interpolation.apply


def offset2batch(offset):
    counts = torch.diff(offset.long(), prepend=offset.new_zeros(1).long())
    return torch.repeat_interleave(torch.arange(len(offset), device=offset.device), counts)


def batch2offset(batch):
    return torch.cumsum(batch.bincount(), dim=0).int()
