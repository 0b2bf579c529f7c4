"""Python handle over an fk_grid (the device tables of one build).

All work is enqueued on torch's current HIP stream; inputs must be HIP
device tensors.  This is plumbing over the C ABI: the construction and the
queries themselves run in libflashknn.so.
"""
from __future__ import annotations

import ctypes
import math

import torch

from . import _lib as F


def _dev(t: torch.Tensor, dtype: torch.dtype, name: str) -> torch.Tensor:
    F.require_device(t, name)
    return t.to(dtype).contiguous()


# candidate batch of the first pass -> fk_query_params.flags
BATCH_FLAGS = {"small": F.FK_QUERY_BATCH_SMALL, "default": 0, "large": F.FK_QUERY_BATCH_LARGE}


def _build_flags(no_sync: bool, columns: bool) -> int:
    return (F.FK_BUILD_NO_SYNC if no_sync else 0) | (F.FK_BUILD_COLUMNS if columns else 0)


class Grid:
    """Cell-sorted support points of one or more clouds on one HIP device."""

    # speed hints for exact queries of the last build_exact_grid
    # (FK_QUERY_PRUNE, FK_QUERY_ONE_LANE, the candidate batch), chosen by the
    # autotuner
    prune_hint = False
    one_lane_hint = False
    batch_hint = "default"

    def __init__(self):
        self._L = F.load()
        h = ctypes.c_void_p()
        F.check(self._L.fk_grid_create(ctypes.byref(h)), "fk_grid_create")
        self._h = h
        self._keep = []          # tensors captured by a graph stay alive
        self._params = {}        # query parameter blocks by their arguments
        self.from_xyz = False
        self.num_points = 0
        self.num_clouds = 0
        self.num_queries = -1

    def hints(self):
        """The autotuned speed hints of the last build_exact_grid as keyword
        arguments of query() (results never depend on them)."""
        return dict(prune=self.prune_hint, one_lane=self.one_lane_hint,
                    batch=self.batch_hint)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            torch.cuda.synchronize()
            self._L.fk_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ------------------------------------------------------------ builds
    def reserve(self, max_points, max_cells, max_clouds, max_queries=0):
        F.check(self._L.fk_grid_reserve(self._h, int(max_points), int(max_cells),
                                        int(max_clouds), int(max_queries)),
                "fk_grid_reserve")

    def build_cells(self, grid_coord, batch_idx=None, xyz=None, num_clouds=None,
                    num_down=0, no_sync=False, columns=False):
        """cell = grid_coord >> num_down, cloud = batch_idx (FlashKNN layout).
        ``columns`` forces the column layout (taken by itself when the dense
        cell boxes exceed the table capacity); results do not change."""
        g = _dev(grid_coord, torch.int64, "grid_coord")
        if g.dim() != 2 or g.shape[1] != 3:
            raise ValueError("grid_coord must be [N, 3]")
        n = g.shape[0]
        b = None
        if batch_idx is not None:
            b = _dev(batch_idx, torch.int64, "batch_idx").reshape(-1)
            if b.shape[0] != n:
                raise ValueError("batch_idx must have one entry per point")
            if num_clouds is None:
                num_clouds = int(b.max().item()) + 1 if n else 1
        if num_clouds is None:
            num_clouds = 1
        x = None
        if xyz is not None:
            x = _dev(xyz, torch.float32, "xyz")
            if x.shape != (n, 3):
                raise ValueError("xyz must be [N, 3] like grid_coord")
        F.check(self._L.fk_grid_build_cells(
            self._h, F.ptr(x), F.ptr(g), F.ptr(b), n, int(num_clouds), int(num_down),
            _build_flags(no_sync, columns), F.stream_ptr()), "fk_grid_build_cells")
        self.from_xyz = False
        self.num_points, self.num_clouds, self.num_queries = n, int(num_clouds), -1
        return self

    def build_xyz(self, xyz, offsets, inv_cell, no_sync=False, columns=False):
        """cell = floor(xyz * inv_cell[c]); clouds by cumulative end offsets."""
        x = _dev(xyz, torch.float32, "xyz")
        if x.dim() != 2 or x.shape[1] != 3:
            raise ValueError("xyz must be [N, 3]")
        o = _dev(offsets, torch.int32, "offsets").reshape(-1)
        inv = _dev(inv_cell, torch.float32, "inv_cell").reshape(-1)
        if inv.shape[0] != o.shape[0]:
            raise ValueError("one inverse cell size per cloud")
        F.check(self._L.fk_grid_build_xyz(
            self._h, F.ptr(x), F.ptr(o), int(o.shape[0]), x.shape[0], F.ptr(inv),
            _build_flags(no_sync, columns), F.stream_ptr()), "fk_grid_build_xyz")
        self.from_xyz = True
        self.num_points, self.num_clouds, self.num_queries = x.shape[0], int(o.shape[0]), -1
        return self

    def auto_cell(self, xyz, offsets, k, scale):
        """(side, inv) float32 [num_clouds] from the bounding-box density.
        ``k`` is an int or an int32 device tensor of per-cloud k."""
        x = _dev(xyz, torch.float32, "xyz")
        o = _dev(offsets, torch.int32, "offsets").reshape(-1)
        B = o.shape[0]
        side = torch.empty(B, dtype=torch.float32, device=x.device)
        inv = torch.empty(B, dtype=torch.float32, device=x.device)
        kt = None
        if isinstance(k, torch.Tensor):
            kt = _dev(k, torch.int32, "k per cloud").reshape(-1)
            kk = 0
        else:
            kk = int(k)
        F.check(self._L.fk_grid_auto_cell(self._h, F.ptr(x), F.ptr(o), B, x.shape[0],
                                          F.ptr(kt), kk, float(scale), F.ptr(side),
                                          F.ptr(inv), F.stream_ptr()), "fk_grid_auto_cell")
        return side, inv

    def adapt_cell(self, k, side, inv, target_ratio=9.0, exponent=2.0):
        """Density feedback on the current xyz build (in place on side/inv).
        ``target_ratio``: a float, or a float32 device tensor per cloud."""
        kt = None
        if isinstance(k, torch.Tensor):
            kt = _dev(k, torch.int32, "k per cloud").reshape(-1)
            kk = 0
        else:
            kk = int(k)
        rt, r = None, 1.0
        if isinstance(target_ratio, torch.Tensor):
            rt = _dev(target_ratio, torch.float32, "target ratios").reshape(-1)
        else:
            r = float(target_ratio)
        F.check(self._L.fk_grid_adapt_cell(self._h, F.ptr(kt), kk, r, F.ptr(rt),
                                           float(exponent), F.ptr(side), F.ptr(inv),
                                           F.stream_ptr()), "fk_grid_adapt_cell")
        return side, inv

    def cloud_candidates(self):
        """int64 [num_clouds]: sum of 27-cell candidate counts of each cloud."""
        out = torch.empty(self.num_clouds, dtype=torch.int64,
                          device=torch.device("cuda", torch.cuda.current_device()))
        F.check(self._L.fk_grid_cloud_candidates(self._h, F.ptr(out), F.stream_ptr()),
                "fk_grid_cloud_candidates")
        return out

    def attach_queries(self, qxyz, qcell=None, qbatch=None, qoffsets=None, qpad=None,
                       no_sync=False):
        q = _dev(qxyz, torch.float32, "query xyz")
        if q.dim() != 2 or q.shape[1] != 3:
            raise ValueError("query xyz must be [M, 3]")
        c = None if qcell is None else _dev(qcell, torch.int64, "query cells")
        b = None if qbatch is None else _dev(qbatch, torch.int64, "query batch").reshape(-1)
        o = None if qoffsets is None else _dev(qoffsets, torch.int32, "query offsets").reshape(-1)
        p = None if qpad is None else _dev(qpad, torch.int32, "query pad").reshape(-1)
        F.check(self._L.fk_grid_attach_queries(
            self._h, F.ptr(q), F.ptr(c), F.ptr(b), F.ptr(o), q.shape[0], F.ptr(p),
            F.FK_BUILD_NO_SYNC if no_sync else 0, F.stream_ptr()), "fk_grid_attach_queries")
        self.num_queries = q.shape[0]
        return self

    def status(self):
        F.check(self._L.fk_grid_status(self._h, F.stream_ptr()), "device status")

    def info(self):
        n, c, cells, q = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64()
        F.check(self._L.fk_grid_info(self._h, ctypes.byref(n), ctypes.byref(c),
                                     ctypes.byref(cells), ctypes.byref(q)), "fk_grid_info")
        return {"points": n.value, "clouds": c.value, "cells": cells.value, "queries": q.value}

    def tables(self):
        """(sorted records float[N,4], cell_start int32[cells+1]) as tensors.

        The tensors are views of the handle's own device tables: they are
        valid until the next build or growing ``reserve`` on this grid
        (which may reallocate the tables) and must not outlive the grid."""
        info = self.info()
        rec, cs = ctypes.c_void_p(), ctypes.c_void_p()
        F.check(self._L.fk_grid_tables(self._h, ctypes.byref(rec), ctypes.byref(cs)),
                "fk_grid_tables")
        n, cells = info["points"], info["cells"]
        dev = torch.device("cuda", torch.cuda.current_device())
        return (_wrap(rec.value, (n, 4), torch.float32, dev),
                _wrap(cs.value, (cells + 1,), torch.int32, dev))

    # ----------------------------------------------------------- queries
    def query(self, k, *, mode=F.FK_MODE_CELLS, radius_cells=1, cut_r2=math.inf,
              idx_dtype=torch.int64, pad_mode=F.FK_PAD_SELF, dist_mode=F.FK_DIST_SQ,
              pad_dist=math.inf, attached=False, out_idx=None, out_dist=None,
              passes=0, prune=False, one_lane=False, uniform=False, batch="default"):
        """passes: 0 = full query; else a mask (1 reset, 2 first pass,
        4 expansion) used to time one kernel alone.  prune / one_lane /
        uniform: FK_QUERY_PRUNE / FK_QUERY_ONE_LANE / FK_QUERY_UNIFORM;
        batch: "small" / "default" / "large" candidate batch
        (FK_QUERY_BATCH_SMALL / _LARGE).  Speed hints: results are identical."""
        if batch not in BATCH_FLAGS:
            raise ValueError(f"batch must be one of {sorted(BATCH_FLAGS)}")
        nq = self.num_queries if attached else self.num_points
        if attached and nq < 0:
            raise RuntimeError("no queries attached")
        dev = torch.device("cuda", torch.cuda.current_device())
        if out_idx is None:
            out_idx = torch.empty((nq, k), dtype=idx_dtype, device=dev)
        if dist_mode != F.FK_DIST_NONE and out_dist is None:
            out_dist = torch.empty((nq, k), dtype=torch.float32, device=dev)
        key = (k, mode, radius_cells, cut_r2, out_idx.dtype == torch.int64, pad_mode,
               dist_mode, pad_dist, attached, passes, prune, one_lane, uniform, batch)
        p = self._params.get(key)
        if p is None:  # the parameter block of a repeated query is built once
            p = F.make_params(k, mode, radius_cells, cut_r2, out_idx.dtype == torch.int64,
                              pad_mode, dist_mode, pad_dist,
                              F.FK_QUERIES_ATTACHED if attached else F.FK_QUERIES_SELF,
                              passes, (F.FK_QUERY_PRUNE if prune else 0)
                              | (F.FK_QUERY_ONE_LANE if one_lane else 0)
                              | (F.FK_QUERY_UNIFORM if uniform else 0) | BATCH_FLAGS[batch])
            if len(self._params) > 64:
                self._params.clear()
            self._params[key] = p
        F.check(self._L.fk_query(self._h, ctypes.byref(p), F.ptr(out_idx),
                                 F.ptr(out_dist), F.stream_ptr()), "fk_query")
        return out_idx, out_dist

    def query_stats(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        F.check(self._L.fk_query_stats(self._h, ctypes.byref(a), ctypes.byref(b),
                                       F.stream_ptr()), "fk_query_stats")
        return {"certified_first_pass": a.value, "expanded": b.value}

    def candidate_count(self, radius_cells=1, attached=False):
        """Sum over queries of |C(q)| for the block of half-width radius_cells."""
        t = ctypes.c_int64()
        F.check(self._L.fk_candidate_count(
            self._h, int(radius_cells),
            F.FK_QUERIES_ATTACHED if attached else F.FK_QUERIES_SELF,
            ctypes.byref(t), F.stream_ptr()), "fk_candidate_count")
        return t.value

    def point_candidates(self, radius_cells=1, attached=False):
        """int32 [nq]: |C(q)| of every query (row order of the query set)."""
        nq = self.num_queries if attached else self.num_points
        out = torch.empty(max(nq, 0), dtype=torch.int32,
                          device=torch.device("cuda", torch.cuda.current_device()))
        F.check(self._L.fk_point_candidates(
            self._h, int(radius_cells),
            F.FK_QUERIES_ATTACHED if attached else F.FK_QUERIES_SELF,
            F.ptr(out), F.stream_ptr()), "fk_point_candidates")
        return out

    def batched_query(self, k_per_cloud, out_offsets, out_idx, out_dist, *, max_k,
                      mode=F.FK_MODE_EXACT, radius_cells=1, idx_dtype=torch.int64,
                      pad_mode=F.FK_PAD_SELF, dist_mode=F.FK_DIST_SQ, capture=False,
                      prune=False):
        p = F.make_params(max_k, mode, radius_cells, math.inf, idx_dtype == torch.int64,
                          pad_mode, dist_mode, math.inf, F.FK_QUERIES_SELF,
                          flags=F.FK_QUERY_PRUNE if prune else 0)
        fn = self._L.fk_batched_capture if capture else self._L.fk_batched_query
        F.check(fn(self._h, ctypes.byref(p), F.ptr(k_per_cloud), F.ptr(out_offsets),
                   F.ptr(out_idx), F.ptr(out_dist), F.stream_ptr()),
                "fk_batched_capture" if capture else "fk_batched_query")
        if capture:
            self._keep = [k_per_cloud, out_offsets, out_idx, out_dist]

    def replay(self):
        F.check(self._L.fk_graph_replay(self._h, F.stream_ptr()), "fk_graph_replay")


def _wrap(addr, shape, dtype, device):
    """Non-owning tensor view of library-owned device memory."""
    numel = 1
    for s in shape:
        numel *= s
    if numel == 0 or not addr:
        return torch.empty(shape, dtype=dtype, device=device)

    class _Iface:
        __cuda_array_interface__ = {
            "shape": tuple(shape),
            "typestr": {torch.float32: "<f4", torch.int32: "<i4"}[dtype],
            "data": (addr, False),
            "version": 3,
            "strides": None,
        }
    return torch.as_tensor(_Iface(), device=device)
