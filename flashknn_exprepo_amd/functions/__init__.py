"""Drop-in for the reference's ``FlashKNN/functions`` package
(FlashKNN/functions/__init__.py:1-8)."""
from .FlashKnnWrapper import FlashKNN
from .adaptive_octree import (
    AdaptiveNeighborhoodFlashKNN,
    build_adaptive_octree,
    select_adaptive_levels,
)
from .utils import cal_recall, vanilla_Knn
from .z_order import neighbor_keys, xyz2key

__all__ = ["FlashKNN", "AdaptiveNeighborhoodFlashKNN", "build_adaptive_octree",
           "select_adaptive_levels", "xyz2key", "neighbor_keys", "cal_recall",
           "vanilla_Knn"]
