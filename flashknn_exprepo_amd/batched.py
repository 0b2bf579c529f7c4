"""Batched kNN over many uneven clouds in one persistent launch.

Clouds are packed by int32 cumulative end offsets (the pointops layout).
One build sorts all clouds into their own dense cell boxes; one persistent
kernel then pulls (cloud, 64-query tile) work items from a device queue,
each tile with its cloud's k (8/16/32/64 register classes), and an
expansion kernel finishes the queries the first pass could not certify.
The query step can be captured once into a hipGraph and replayed.

No reference counterpart: the reference launches one FlashKNN query per
cloud (FlashKNN/functions/FlashKnnWrapper.py:165) with a single K.
"""
from __future__ import annotations

import torch

from . import _lib as F
from .grid import Grid
from .knn import build_exact_grid


class BatchedKNN:
    """Exact (or 27-cell) kNN of every point of every cloud, per-cloud k."""

    def __init__(self, xyz, offsets, k_per_cloud, *, exact=True, cell=None,
                 scale=None, target_ratio=None, idx_dtype=torch.int64):
        F.require_device(xyz, "xyz")
        self.xyz = xyz.float().contiguous()
        self.offsets = offsets.to(device=xyz.device, dtype=torch.int32).contiguous()
        ks = torch.as_tensor(k_per_cloud, dtype=torch.int32).to(xyz.device).contiguous()
        if ks.shape[0] != self.offsets.shape[0]:
            raise ValueError("one k per cloud")
        kmin, kmax = int(ks.min()), int(ks.max())
        if kmin < 1 or kmax > F.FK_MAX_BATCHED_K:
            raise ValueError(f"k must be in [1, {F.FK_MAX_BATCHED_K}]")
        self.k = ks
        self.max_k = kmax
        self.exact = exact
        ends = self.offsets.long()
        starts = torch.cat([ends.new_zeros(1), ends[:-1]])
        sizes = ends - starts
        rows = sizes * ks.long()
        self.out_offsets = torch.cumsum(rows, 0) - rows          # element offsets
        total = int(rows.sum())
        self.sizes = sizes
        self.out_idx = torch.empty(total, dtype=idx_dtype, device=xyz.device)
        self.out_dist = torch.empty(total, dtype=torch.float32, device=xyz.device)
        self.grid = Grid()
        # per-cloud first guess + density feedback (sizes read back: sync builds)
        self.cell = build_exact_grid(self.grid, self.xyz, self.offsets, self.k, cell=cell,
                                     scale=scale, target_ratio=target_ratio,
                                     no_sync=False)
        self._captured = False

    def run(self):
        """One batched query step (eager launch)."""
        self.grid.batched_query(self.k, self.out_offsets, self.out_idx, self.out_dist,
                                max_k=self.max_k,
                                mode=F.FK_MODE_EXACT if self.exact else F.FK_MODE_CELLS,
                                idx_dtype=self.out_idx.dtype, prune=self.grid.prune_hint)
        return self.out_idx, self.out_dist

    def capture(self, stream=None):
        """Capture the query step into the handle's hipGraph (needs a
        non-default stream; default priority)."""
        s = stream or torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.grid.batched_query(self.k, self.out_offsets, self.out_idx,
                                    self.out_dist, max_k=self.max_k,
                                    mode=F.FK_MODE_EXACT if self.exact else F.FK_MODE_CELLS,
                                    idx_dtype=self.out_idx.dtype, capture=True,
                                    prune=self.grid.prune_hint)
        torch.cuda.current_stream().wait_stream(s)
        self._captured = True

    def replay(self):
        if not self._captured:
            self.capture()
        self.grid.replay()
        return self.out_idx, self.out_dist

    def per_cloud(self, c):
        """(idx [n_c, k_c], dist [n_c, k_c]) views of cloud c's rows."""
        o = int(self.out_offsets[c])
        n, k = int(self.sizes[c]), int(self.k[c])
        return (self.out_idx[o:o + n * k].view(n, k),
                self.out_dist[o:o + n * k].view(n, k))
