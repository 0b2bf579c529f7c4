"""Drop-in for the reference's ``FlashKNN/functions`` package
(FlashKNN/functions/__init__.py:1-8)."""
from .FlashKnnWrapper import FlashKNN
from .utils import cal_recall, vanilla_Knn
from .z_order import neighbor_keys, xyz2key

__all__ = ["FlashKNN", "xyz2key", "neighbor_keys", "cal_recall", "vanilla_Knn"]
