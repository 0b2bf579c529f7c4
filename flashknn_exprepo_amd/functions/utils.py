"""Brute-force kNN and multiset recall (FlashKNN/functions/utils.py:5-104).

``vanilla_Knn(query, support, K)`` keeps the reference's contract: a blocked
brute force by the matmul expansion ||q||^2 + ||s||^2 - 2 q.s with a running
top-K, returning (L2 distance [M, K], index [M, K]).  ``cal_recall`` is the
multiset recall per row (duplicates are counted min(pred, gt) times).
These are the reference's own helpers; they run with torch on whatever
device the inputs live on and are not part of the kernel path.
"""
from __future__ import annotations

import torch


def vanilla_Knn(query: torch.Tensor, support: torch.Tensor, K: int,
                query_batchsize: int = 1000, support_batchsize: int = 10000):
    if query.device != support.device:
        raise AssertionError("query and support must be on the same device")
    if query.size(1) != support.size(1):
        raise AssertionError("Feature dimensions must match")
    M, N = query.shape[0], support.shape[0]
    dev = query.device
    out_d = torch.full((M, K), float("inf"), device=dev)
    out_i = torch.full((M, K), -1, device=dev, dtype=torch.long)
    s_sq = (support * support).sum(1)
    for q0 in range(0, M, query_batchsize):
        q = query[q0:q0 + query_batchsize]
        q_sq = (q * q).sum(1, keepdim=True)
        best_d = torch.full((q.shape[0], K), float("inf"), device=dev)
        best_i = torch.full((q.shape[0], K), -1, device=dev, dtype=torch.long)
        for s0 in range(0, N, support_batchsize):
            s = support[s0:s0 + support_batchsize]
            d = q_sq + s_sq[s0:s0 + s.shape[0]][None, :] - 2.0 * (q @ s.t())
            cand_d = torch.cat([best_d, d], 1)
            ids = torch.arange(s0, s0 + s.shape[0], device=dev).expand(q.shape[0], -1)
            cand_i = torch.cat([best_i, ids], 1)
            kk = min(K, cand_d.shape[1])
            best_d, pos = torch.topk(cand_d, kk, dim=1, largest=False)
            best_i = cand_i.gather(1, pos)
        out_d[q0:q0 + q.shape[0], :best_d.shape[1]] = torch.sqrt(best_d.clamp_min(0))
        out_i[q0:q0 + q.shape[0], :best_i.shape[1]] = best_i
    return out_d, out_i


def cal_recall(nbr_indices_pred: torch.Tensor, nbr_indices_gt: torch.Tensor) -> torch.Tensor:
    """Per-row multiset recall: sum over values of min(count_pred, count_gt) / K."""
    pred = nbr_indices_pred.long()
    gt = nbr_indices_gt.long()
    K = pred.shape[1]
    allv = torch.cat([gt, pred], 1)
    sorted_v, _ = allv.sort(1)
    # rank of each value inside its row (distinct values get distinct slots)
    new = torch.ones_like(sorted_v, dtype=torch.bool)
    new[:, 1:] = sorted_v[:, 1:] != sorted_v[:, :-1]
    slot_of_sorted = new.long().cumsum(1) - 1
    def slots(v):
        pos = torch.searchsorted(sorted_v, v.contiguous())
        return slot_of_sorted.gather(1, pos)
    cnt_gt = torch.zeros_like(sorted_v).scatter_add_(1, slots(gt), torch.ones_like(gt))
    cnt_pr = torch.zeros_like(sorted_v).scatter_add_(1, slots(pred), torch.ones_like(pred))
    return torch.minimum(cnt_gt, cnt_pr).sum(1).to(torch.float32) / K
