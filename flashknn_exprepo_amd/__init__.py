"""MI355X-native FlashKNN: grid kNN for point clouds on gfx950.

The compute path is libflashknn.so (hand-written HIP kernels behind a C
ABI, include/flashknn.h), loaded with ctypes; torch supplies device memory
and streams.  Public surface:

* ``FlashKNN``, ``xyz2key``, ``cal_recall``, ``vanilla_Knn`` - drop-in for
  the reference's ``FlashKNN.functions``;
* ``pointops.knn_query`` / ``interpolation`` / ``interpolation2`` - the
  offset-batched layout of Pointcept's pointops;
* ``knn`` / ``ExactKNN`` - certified exact kNN on the grid;
* ``BatchedKNN`` - many uneven clouds, per-cloud k, one persistent launch,
  hipGraph replay;
* ``gathered_sqdist`` / ``idw_interpolate`` - differentiable distances and
  inverse-distance-weighted interpolation with hand-written backward;
* ``voxel_first`` / ``build_flash_hierarchy`` - voxel-first subsampling on
  the device and the level hierarchy built from it.
"""
from . import pointops
from ._lib import FlashKNNError, load
from .autograd import gathered_sqdist, idw_gather, idw_interpolate
from .batched import BatchedKNN
from .functions import (AdaptiveNeighborhoodFlashKNN, FlashKNN, build_adaptive_octree,
                        cal_recall, select_adaptive_levels, vanilla_Knn, xyz2key)
from .grid import Grid
from .hierarchy import build_flash_hierarchy, voxel_first
from .knn import ExactKNN, StaticKNN, cell_sizes, knn

__all__ = [
    "FlashKNN", "AdaptiveNeighborhoodFlashKNN", "build_adaptive_octree",
    "select_adaptive_levels", "xyz2key", "cal_recall", "vanilla_Knn", "pointops", "knn",
    "ExactKNN", "StaticKNN", "cell_sizes", "BatchedKNN", "Grid", "gathered_sqdist",
    "idw_gather", "idw_interpolate", "FlashKNNError", "load", "voxel_first",
    "build_flash_hierarchy",
]
