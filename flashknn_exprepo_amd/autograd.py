"""Differentiable gathered distances and IDW interpolation.

* ``GatheredSqDist``: d_ij = ||q_i - p_idx(i,j)||^2 with the neighbour
  indices treated as constants.  Backward (libflashknn fk_dist_backward):
  dL/dq_i = sum_j 2 g_ij (q_i - p_j); dL/dp_m = -sum_{idx=m} 2 g_ij (q_i - p_m),
  the second a deterministic segmented sum over a reverse CSR of idx.
* ``IDWGather``: out_i = sum_j w_ij feat[idx_ij] (fk_idw_forward) with
  backward grad_feat (segmented, deterministic) and grad_w (per entry).
* ``idw_interpolate``: weights w = 1/(d + eps) on the squared distance,
  normalised per row, as one fused forward (fk_idw_interp_forward:
  distances, weights and gather in one wave per row) and one fused backward
  (fk_idw_interp_backward: gradients for features, queries and supports
  over a single reverse CSR).  Unlike pointops' 1/(sqrt(d) + 1e-8)
  (Pointcept/libs/pointops/functions/interpolation.py:15-17), whose gradient
  is unbounded at d = 0 (coarse points are a subset of the fine ones, so
  self matches have d = 0), this form is smooth there.

The reference differentiates no coordinates (FlashKNN.query is
@torch.no_grad, FlashKnnWrapper.py:165; pointops returns None for xyz,
interpolation.py:56); this layer is new.
"""
from __future__ import annotations

import torch
from torch.autograd import Function

from . import _lib as F


def _ws(nq, ns, k, device):
    L = F.load()
    nbytes = int(L.fk_backward_workspace_bytes(int(nq), int(ns), int(k)))
    return torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)


def gathered_sqdist_forward(q, p, idx):
    """d_ij in fp32 as ((dx*dx + dy*dy) + dz*dz) (torch elementwise ops)."""
    safe = idx.clamp_min(0)
    diff = q[:, None, :] - p[safe]
    d = (diff[..., 0] * diff[..., 0] + diff[..., 1] * diff[..., 1]) + diff[..., 2] * diff[..., 2]
    return torch.where(idx >= 0, d, torch.full_like(d, float("inf")))


class GatheredSqDist(Function):
    @staticmethod
    def forward(ctx, q, p, idx, d=None):
        q = q.float().contiguous()
        p = p.float().contiguous()
        idx = idx.long().contiguous()
        d = gathered_sqdist_forward(q, p, idx) if d is None else d.float().clone()
        ctx.save_for_backward(q, p, idx)
        ctx.mark_non_differentiable(idx)
        return d

    @staticmethod
    def backward(ctx, grad_d):
        q, p, idx = ctx.saved_tensors
        g = grad_d.float().contiguous()
        g = torch.where(idx >= 0, g, torch.zeros_like(g))
        nq, k = idx.shape
        ns = p.shape[0]
        need_q, need_p = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        gq = torch.empty_like(q) if need_q else None
        gp = torch.empty_like(p) if need_p else None
        ws = _ws(nq, ns, k, q.device) if need_p else None
        L = F.load()
        F.check(L.fk_dist_backward(F.ptr(q), nq, F.ptr(p), ns, F.ptr(idx), k, F.ptr(g),
                                   F.ptr(gq), F.ptr(gp), F.ptr(ws), F.stream_ptr()),
                "fk_dist_backward")
        return gq, gp, None, None


gathered_sqdist = GatheredSqDist.apply


def idw_gather_forward(feat, idx, w):
    nq, k = idx.shape
    out = torch.empty(nq, feat.shape[1], dtype=torch.float32, device=feat.device)
    L = F.load()
    F.check(L.fk_idw_forward(F.ptr(feat), feat.shape[0], feat.shape[1], F.ptr(idx),
                             F.ptr(w), nq, k, F.ptr(out), F.stream_ptr()),
            "fk_idw_forward")
    return out


def idw_gather_backward(feat, ns, idx, w, grad_out, need_w=True, need_feat=True):
    nq, k = idx.shape
    C = grad_out.shape[1]
    dev = grad_out.device
    gf = torch.empty(ns, C, dtype=torch.float32, device=dev) if need_feat else None
    gw = torch.empty(nq, k, dtype=torch.float32, device=dev) if need_w else None
    ws = _ws(nq, ns, k, dev) if need_feat else None
    L = F.load()
    F.check(L.fk_idw_backward(F.ptr(feat), ns, C, F.ptr(idx), F.ptr(w), nq, k,
                              F.ptr(grad_out), F.ptr(gf), F.ptr(gw), F.ptr(ws),
                              F.stream_ptr()), "fk_idw_backward")
    return gf, gw


class IDWGather(Function):
    @staticmethod
    def forward(ctx, feat, w, idx):
        feat = feat.float().contiguous()
        w = w.float().contiguous()
        idx = idx.long().contiguous()
        out = idw_gather_forward(feat, idx, w)
        ctx.save_for_backward(feat, w, idx)
        ctx.mark_non_differentiable(idx)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        feat, w, idx = ctx.saved_tensors
        gf, gw = idw_gather_backward(feat, feat.shape[0], idx, w,
                                     grad_out.float().contiguous(),
                                     need_w=ctx.needs_input_grad[1],
                                     need_feat=ctx.needs_input_grad[0])
        return gf, gw, None


idw_gather = IDWGather.apply


class IDWInterp(Function):
    """out_i = sum_j w_ij feat[idx_ij], w_ij = r_ij / sum_l r_il,
    r_ij = 1 / (||p_idx - q_i||^2 + eps) (0 for idx < 0)."""

    @staticmethod
    def forward(ctx, feat, q, p, idx, eps):
        feat = feat.float().contiguous()
        q = q.float().contiguous()
        p = p.float().contiguous()
        idx = idx.long().contiguous()
        nq, k = idx.shape
        if k > 64:
            raise ValueError("idw_interpolate supports k <= 64")
        out = torch.empty(nq, feat.shape[1], dtype=torch.float32, device=feat.device)
        L = F.load()
        F.check(L.fk_idw_interp_forward(F.ptr(q), nq, F.ptr(p), p.shape[0], F.ptr(idx), k,
                                        F.ptr(feat), feat.shape[1], float(eps), None,
                                        F.ptr(out), F.stream_ptr()), "fk_idw_interp_forward")
        ctx.save_for_backward(feat, q, p, idx)
        ctx.eps = float(eps)
        ctx.mark_non_differentiable(idx)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        feat, q, p, idx = ctx.saved_tensors
        nq, k = idx.shape
        ns, C = feat.shape
        need_f, need_q, need_p = ctx.needs_input_grad[:3]
        gf = torch.empty_like(feat) if need_f else None
        gq = torch.empty_like(q) if need_q else None
        gp = torch.empty_like(p) if need_p else None
        L = F.load()
        nbytes = int(L.fk_idw_interp_workspace_bytes(nq, ns, k))
        ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=feat.device)
        g = grad_out.float().contiguous()
        F.check(L.fk_idw_interp_backward(F.ptr(q), nq, F.ptr(p), ns, F.ptr(idx), k,
                                         F.ptr(feat), C, ctx.eps, F.ptr(g), F.ptr(gf),
                                         F.ptr(gq), F.ptr(gp), F.ptr(ws), F.stream_ptr()),
                "fk_idw_interp_backward")
        return gf, gq, gp, None, None


def idw_interpolate(feat, q, p, idx, eps=1e-8, d=None):
    """Interpolate feat (on support p) at queries q from their neighbours idx.

    Differentiable w.r.t. feat, q and p (idx constant).  The distances are
    recomputed in the query kernels' fp32 form, so a ``d`` returned by the
    kNN query equals them and is not needed (accepted for compatibility)."""
    return IDWInterp.apply(feat, q, p, idx, eps)
