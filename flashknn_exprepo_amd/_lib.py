"""ctypes binding of libflashknn.so (the C ABI declared in include/flashknn.h).

The library is loaded after torch so that its libamdhip64.so.7 dependency
resolves to the HIP runtime torch already loaded (same SONAME).  There is
no CPU fallback: if the library is missing or no HIP device is present,
every entry point raises.
"""
from __future__ import annotations

import ctypes
import math
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflashknn.so")

FK_OK = 0
FK_ERR_INVALID_ARG = 1
FK_ERR_K_RANGE = 2
FK_ERR_GRID_TOO_LARGE = 3
FK_ERR_DOMAIN = 4
FK_ERR_HIP = 5
FK_ERR_CAPACITY = 6
FK_ERR_STATE = 7
FK_MAX_K = 128
FK_MAX_BATCHED_K = 64

FK_MODE_CELLS = 0
FK_MODE_EXACT = 1
FK_PAD_SELF = 0
FK_PAD_MINUS1 = 1
FK_PAD_PER_QUERY = 2
FK_DIST_NONE = 0
FK_DIST_SQ = 1
FK_DIST_SQRT = 2
FK_QUERIES_SELF = 0
FK_QUERIES_ATTACHED = 1
FK_BUILD_NO_SYNC = 1
FK_BUILD_COLUMNS = 2
FK_QUERY_PRUNE = 1
FK_QUERY_ONE_LANE = 2
FK_QUERY_UNIFORM = 4
FK_QUERY_BATCH_SMALL = 8
FK_QUERY_BATCH_LARGE = 16
FK_VOXEL_NOALLOC_MAX_POINTS = 1048576
FK_ADAPTIVE_STATS = 72


class FlashKNNError(RuntimeError):
    """A libflashknn call failed; ``status`` holds the fk_status code."""

    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


class QueryParams(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("radius_cells", ctypes.c_int32),
        ("cut_r2", ctypes.c_float),
        ("idx_64", ctypes.c_int32),
        ("pad_mode", ctypes.c_int32),
        ("dist_mode", ctypes.c_int32),
        ("pad_dist", ctypes.c_float),
        ("query_set", ctypes.c_int32),
        ("passes", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


def make_params(k, mode=FK_MODE_CELLS, radius_cells=1, cut_r2=math.inf, idx_64=True,
                pad_mode=FK_PAD_SELF, dist_mode=FK_DIST_SQ, pad_dist=math.inf,
                query_set=FK_QUERIES_SELF, passes=0, flags=0) -> QueryParams:
    p = QueryParams()
    p.k = int(k)
    p.mode = int(mode)
    p.radius_cells = int(radius_cells)
    p.cut_r2 = float(cut_r2)
    p.idx_64 = 1 if idx_64 else 0
    p.pad_mode = int(pad_mode)
    p.dist_mode = int(dist_mode)
    p.pad_dist = float(pad_dist)
    p.query_set = int(query_set)
    p.passes = int(passes)
    p.flags = int(flags)
    return p


# (name, restype, argtypes) of every entry point in include/flashknn.h
_vp = ctypes.c_void_p
_i32, _i64 = ctypes.c_int32, ctypes.c_int64
SIGNATURES = [
    ("fk_status_string", ctypes.c_char_p, [ctypes.c_int]),
    ("fk_last_error", ctypes.c_char_p, []),
    ("fk_abi_version", ctypes.c_int, []),
    ("fk_morton_encode", ctypes.c_int, [_vp, _vp, _i64, _vp, _vp]),
    ("fk_morton_neighbor_keys", ctypes.c_int, [_vp, _vp, _i64, _vp, _vp]),
    ("fk_grid_create", ctypes.c_int, [ctypes.POINTER(_vp)]),
    ("fk_grid_destroy", ctypes.c_int, [_vp]),
    ("fk_grid_reserve", ctypes.c_int, [_vp, _i64, _i64, _i32, _i64]),
    ("fk_grid_build_cells", ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _vp]),
    ("fk_grid_build_xyz", ctypes.c_int, [_vp, _vp, _vp, _i32, _i64, _vp, _i32, _vp]),
    ("fk_grid_attach_queries", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, _vp, _i32, _vp]),
    ("fk_grid_auto_cell", ctypes.c_int, [_vp, _vp, _vp, _i32, _i64, _vp, _i32, ctypes.c_float,
                                         _vp, _vp, _vp]),
    ("fk_grid_adapt_cell", ctypes.c_int, [_vp, _vp, _i32, ctypes.c_float, _vp,
                                          ctypes.c_float, _vp, _vp, _vp]),
    ("fk_grid_cloud_candidates", ctypes.c_int, [_vp, _vp, _vp]),
    ("fk_grid_status", ctypes.c_int, [_vp, _vp]),
    ("fk_grid_info", ctypes.c_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i32),
                                    ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    ("fk_grid_tables", ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    ("fk_query", ctypes.c_int, [_vp, ctypes.POINTER(QueryParams), _vp, _vp, _vp]),
    ("fk_query_stats", ctypes.c_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), _vp]),
    ("fk_candidate_count", ctypes.c_int, [_vp, _i32, _i32, ctypes.POINTER(_i64), _vp]),
    ("fk_point_candidates", ctypes.c_int, [_vp, _i32, _i32, _vp, _vp]),
    ("fk_batched_query", ctypes.c_int, [_vp, ctypes.POINTER(QueryParams), _vp, _vp, _vp, _vp, _vp]),
    ("fk_batched_capture", ctypes.c_int, [_vp, ctypes.POINTER(QueryParams), _vp, _vp, _vp, _vp, _vp]),
    ("fk_graph_replay", ctypes.c_int, [_vp, _vp]),
    ("fk_idw_interp_workspace_bytes", ctypes.c_int64, [_i64, _i64, _i32]),
    ("fk_idw_interp_forward", ctypes.c_int, [_vp, _i64, _vp, _i64, _vp, _i32, _vp, _i32,
                                             ctypes.c_float, _vp, _vp, _vp]),
    ("fk_idw_interp_backward", ctypes.c_int, [_vp, _i64, _vp, _i64, _vp, _i32, _vp, _i32,
                                              ctypes.c_float, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("fk_adaptive_workspace_bytes", ctypes.c_int64, [_i64]),
    ("fk_adaptive_select", ctypes.c_int, [_vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _vp,
                                          _vp, _vp, _vp]),
    ("fk_adaptive_query", ctypes.c_int, [_vp, _i64, _i32, ctypes.c_float, _vp, _vp, _vp, _vp]),
    ("fk_voxel_workspace_bytes", ctypes.c_int64, [_i64]),
    ("fk_voxel_first", ctypes.c_int, [_vp, _vp, _i32, _i64, ctypes.c_float, _vp, _vp, _vp, _vp]),
    ("fk_voxel_first_noalloc", ctypes.c_int, [_vp, _vp, _i32, _i64, ctypes.c_float, _vp, _vp, _vp]),
    ("fk_backward_workspace_bytes", ctypes.c_int64, [_i64, _i64, _i32]),
    ("fk_dist_backward", ctypes.c_int, [_vp, _i64, _vp, _i64, _vp, _i32, _vp, _vp, _vp, _vp, _vp]),
    ("fk_idw_forward", ctypes.c_int, [_vp, _i64, _i32, _vp, _vp, _i64, _i32, _vp, _vp]),
    ("fk_idw_backward", ctypes.c_int, [_vp, _i64, _i32, _vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp]),
]

_lib = None


def load(path: str = None):
    """Load and type libflashknn.so (raises if it was never built).

    FLASHKNN_LIB overrides the path (used to compare builds of the same
    sources with different tuning knobs)."""
    global _lib
    if _lib is not None:
        return _lib
    if path is None:
        path = os.environ.get("FLASHKNN_LIB", LIB_PATH)
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int, what: str = "") -> None:
    if status == FK_OK:
        return
    L = load()
    msg = L.fk_last_error().decode() or L.fk_status_string(status).decode()
    text = f"{what}: {msg}" if what else msg
    if status in (FK_ERR_INVALID_ARG, FK_ERR_K_RANGE, FK_ERR_DOMAIN):
        raise ValueError(text)
    raise FlashKNNError(status, text)


def require_device(t: torch.Tensor, name: str) -> None:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a HIP device tensor (FlashKNN has no CPU path)")


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


# the current stream's raw handle without building a Stream object (a few
# microseconds per call on the query path)
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_ptr(stream=None):
    if stream is None and _raw_stream is not None:
        return ctypes.c_void_p(_raw_stream(torch.cuda.current_device()))
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)
