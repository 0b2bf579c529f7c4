"""FlashKNN drop-in (FlashKNN/functions/FlashKnnWrapper.py:22-849).

Same class, constructor and method signatures as the reference, backed by
libflashknn.so:

* ``query(grid_coord, batch_idx, xyz, ...)`` (reference :165-449): every
  point queries the support points of its batch whose parent cell
  ``grid_coord >> num_down`` differs by at most one per axis (27 cells);
  returns int64 neighbour indices [N, K] in caller order, rows ascending by
  squared distance.  Missing slots repeat the query's own index.
* ``selected_query(...)`` (:646-849): only ``query_indices`` query; padding
  is the query's own index.
* ``back_query(...)`` (:460-644): every point of ``xyz`` queries
  ``xyz[sel_indices]`` on cells ``(xyz / query_grid_size).long()``; returns
  positions in ``sel_indices`` ([N] for K=1, else ([N, K], dist)).

Differences from the reference, all deliberate and documented: ties between
equal distances go to the smaller index (the reference keeps the first one
its scan order met); memory_mode / sorting_mode / candidate_mode /
enable_skip / thread_group_size / dynamic_load / batch_for_prune are
validated exactly as the reference validates them and then select the same
kernel (one wave64 design replaces the SM/GM/PS/SS ablations), so results
are identical across modes and always sorted; K ranges over 1..128 (above
64 one wave per query);
back_query with K > 1 returns sorted rows padded with (-1, inf).
"""
from __future__ import annotations

import math
import time

import torch

from .. import _lib as F
from ..grid import Grid
from .z_order import xyz2key


class FlashKNN:
    def __init__(self, num_nbr=32, num_down=2, debug=False, print_time=False,
                 device="cuda", multilevel_neigh=False, threshold=256) -> None:
        assert device == "cuda", "FlashKNN runs on a HIP device only"
        self.num_nbr = num_nbr
        self.num_down = num_down
        self.debug = debug
        self.print_time = print_time
        self.threshold = threshold
        self.multilevel_neigh = multilevel_neigh
        self.CudaShareMem = _lds_per_block()
        if self.debug:
            self.time_list = []
        self._grid = None
        self.last_distances = None

    def _grid_handle(self):
        if self._grid is None:
            self._grid = Grid()
        return self._grid

    @staticmethod
    def rng_grid(rng, device="cuda"):
        """Mesh grid of rng^3 in ij order (reference :451-458)."""
        if isinstance(rng, list):
            rng = torch.tensor(rng, dtype=torch.long, device=device)
        g = torch.meshgrid(rng, rng, rng, indexing="ij")
        return torch.stack(g, dim=-1).view(-1, 3)

    # ------------------------------------------------------------- query
    @torch.no_grad()
    def query(self, grid_coord: torch.Tensor, batch_idx: torch.Tensor,
              xyz: torch.Tensor = None, dynamic_load=True, cut_radius=math.inf,
              batch_for_prune=1, memory_mode="SM", sorting_mode="PS",
              candidate_mode="register", enable_skip=True, thread_group_size=None,
              traverse_info=None):
        _validate_modes(memory_mode, sorting_mode, candidate_mode, enable_skip,
                        thread_group_size)
        ts0 = time.time()
        g = self._grid_handle()
        g.build_cells(grid_coord, batch_idx, xyz, num_down=self.num_down)
        if self.debug:
            torch.cuda.synchronize()
            t_build = time.time() - ts0
            if self.print_time:
                print(f"construction: {t_build}")
        ts = time.time()
        idx, dist = g.query(self.num_nbr, mode=F.FK_MODE_CELLS, radius_cells=1,
                            cut_r2=float(cut_radius) ** 2, idx_dtype=torch.int64,
                            pad_mode=F.FK_PAD_SELF, dist_mode=F.FK_DIST_SQ)
        self.last_distances = dist
        if traverse_info is not None:
            traverse_info[0] = mean_candidates(g)
        if self.debug:
            torch.cuda.synchronize()
            t_query = time.time() - ts
            if self.print_time:
                print(f"query: {t_query}")
            self.time_list.append({
                "预处理耗时": t_build,
                "查询耗时": t_query,
                "查询类型": "query",
                "memory_mode": memory_mode,
                "sorting_mode": sorting_mode,
                "candidate_mode": candidate_mode,
                "enable_skip": enable_skip,
                "thread_group_size": ("adaptive" if thread_group_size is None
                                      else thread_group_size),
                "dynamic_load": dynamic_load})
        return idx

    # ---------------------------------------------------- selected_query
    @torch.no_grad()
    def selected_query(self, xyz: torch.Tensor, grid_coord: torch.Tensor,
                       query_indices: torch.Tensor, batch_idx: torch.Tensor,
                       dynamic_load=True, cut_radius=math.inf, batch_for_prune=1,
                       memory_mode="SM"):
        ts0 = time.time()
        g = self._grid_handle()
        g.build_cells(grid_coord, batch_idx, xyz, num_down=self.num_down)
        qi = query_indices.long()
        qsrc = grid_coord.float() if xyz is None else xyz.float()
        qb = None if batch_idx is None else batch_idx[qi]
        g.attach_queries(qsrc[qi], qcell=grid_coord[qi], qbatch=qb,
                         qpad=qi.to(torch.int32))
        t_build = time.time() - ts0
        ts = time.time()
        idx, dist = g.query(self.num_nbr, mode=F.FK_MODE_CELLS, radius_cells=1,
                            cut_r2=float(cut_radius) ** 2, pad_mode=F.FK_PAD_PER_QUERY,
                            dist_mode=F.FK_DIST_SQ, attached=True)
        self.last_distances = dist
        if self.debug:
            torch.cuda.synchronize()
            self.time_list.append({"预处理耗时": t_build, "查询耗时": time.time() - ts,
                                   "查询类型": "selected_query",
                                   "memory_mode": memory_mode,
                                   "dynamic_load": dynamic_load})
        return idx

    # -------------------------------------------------------- back_query
    @torch.no_grad()
    def back_query(self, xyz: torch.Tensor, sel_indices: torch.Tensor,
                   query_grid_size: float, down_grid_size: float,
                   batch_idx: torch.Tensor, dynamic_load=True, memory_mode="SM"):
        F.require_device(xyz, "xyz")
        grid_coord = quotient32(xyz, query_grid_size).long()  # FlashKnnWrapper.py:480
        sel = sel_indices.long()
        g = self._grid_handle()
        sb = None if batch_idx is None else batch_idx[sel]
        nb = None
        if batch_idx is not None:
            nb = int(batch_idx.max().item()) + 1 if batch_idx.numel() else 1
        g.build_cells(grid_coord[sel], sb, xyz[sel], num_clouds=nb, num_down=0)
        g.attach_queries(xyz, qcell=grid_coord, qbatch=batch_idx)
        idx, dist = g.query(self.num_nbr, mode=F.FK_MODE_CELLS, radius_cells=1,
                            pad_mode=F.FK_PAD_MINUS1, dist_mode=F.FK_DIST_SQ,
                            attached=True)
        if bool((idx[:, 0] < 0).any()):
            raise AssertionError("some query points have no support point in their "
                                 "neighbourhood; consider a larger query_grid_size")
        if self.num_nbr == 1:
            return idx.view(-1)
        return idx, dist


def quotient32(x: torch.Tensor, size: float) -> torch.Tensor:
    """fl32(x / fl32(size)): the correctly rounded fp32 quotient of
    ``x / size`` for an fp32 tensor and a Python scalar.  Evaluated in
    float64 and rounded once to fp32, which gives the correctly rounded
    result for a quotient of fp32 values (53 >= 2 * 24 + 2 bits), so the
    cells do not depend on how a torch build implements fp32 division."""
    s = float(torch.tensor(size, dtype=torch.float32))
    return (x.double() / s).float()


def _validate_modes(memory_mode, sorting_mode, candidate_mode, enable_skip,
                    thread_group_size):
    """The reference's argument checks (FlashKnnWrapper.py:196-219)."""
    if candidate_mode not in {"register", "shared"}:
        raise ValueError(f"candidate_mode must be 'register' or 'shared', got {candidate_mode!r}")
    if thread_group_size not in {None, 8, 16, 32}:
        raise ValueError("thread_group_size must be None, 8, 16, or 32, got "
                         f"{thread_group_size!r}")
    if (candidate_mode != "register" or not enable_skip) and not (
            memory_mode == "SM" and sorting_mode == "PS"):
        raise ValueError("candidate placement/skip ablations require memory_mode='SM' "
                         "and sorting_mode='PS'")
    if thread_group_size is not None and not (
            memory_mode == "SM" and sorting_mode == "PS"
            and candidate_mode == "register" and enable_skip):
        raise ValueError("fixed thread-group ablations require memory_mode='SM', "
                         "sorting_mode='PS', register candidates, and skip enabled")
    if memory_mode not in {"SM", "GM"} or sorting_mode not in {"PS", "SS"}:
        raise ValueError("memory_mode must be 'SM' or 'GM' and sorting_mode 'PS' or 'SS'")


def mean_candidates(g: Grid) -> float:
    """Mean 27-cell candidate count over the support points (the statistic the
    reference reports through traverse_info, FlashKnnWrapper.py:351-352)."""
    if g.num_points == 0:
        return float("nan")
    return g.candidate_count(1) / g.num_points


def _lds_per_block() -> int:
    try:
        p = torch.cuda.get_device_properties(torch.cuda.current_device())
        return int(getattr(p, "shared_memory_per_block", 65536))
    except Exception:
        return 65536


__all__ = ["FlashKNN", "xyz2key"]
