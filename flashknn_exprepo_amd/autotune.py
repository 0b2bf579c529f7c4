"""Autotuner for the grid cell size of exact queries.

The kernel's cost per query is (candidates scanned) + (merges) + (the share
of queries the first pass cannot certify, which pay a shell expansion).
Larger cells certify more queries but scan more candidates.  The cell side
is set by density feedback to give ``target_ratio * k`` candidates per
query in the 27-cell block; the best ratio depends on the cloud size, k and
how the points fill space.  ``tune`` times the exact query for a few ratios
on the actual workload and records the fastest per (log2 N, k class)
bucket in a JSON table; ``lookup_target_ratio`` reads it (falling back to
the default).
"""
from __future__ import annotations

import json
import math
import os
import threading

import torch

_LOCK = threading.Lock()
_TABLE = None
TABLE_PATH = os.environ.get(
    "FLASHKNN_TUNE_TABLE",
    os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_table.json"))
RATIOS = (3.0, 4.5, 6.0, 8.0, 11.0)


def _kclass(k: int) -> int:
    return 8 if k <= 8 else 16 if k <= 16 else 32 if k <= 32 else 64


def bucket(n: int, k: int) -> str:
    return f"n{max(0, int(math.log2(max(n, 1))))}_k{_kclass(k)}"


def _load():
    global _TABLE
    if _TABLE is None:
        try:
            with open(TABLE_PATH) as f:
                _TABLE = json.load(f)
        except (OSError, ValueError):
            _TABLE = {}
    return _TABLE


def lookup_target_ratio(n: int, k: int) -> float:
    from .knn import DEFAULT_TARGET_RATIO
    with _LOCK:
        t = _load()
    e = t.get(bucket(n, k))
    return float(e["target_ratio"]) if e else DEFAULT_TARGET_RATIO


def _time_ms(fn, reps=5):
    fn()
    torch.cuda.synchronize()
    st = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    st[0].record()
    for i in range(reps):
        fn()
        st[i + 1].record()
    torch.cuda.synchronize()
    return sorted(st[i].elapsed_time(st[i + 1]) for i in range(reps))[reps // 2]


def tune(xyz, offsets, k, ratios=RATIOS, save=True):
    """Time the exact query at each target ratio; record and return the best."""
    from .knn import ExactKNN
    eng = ExactKNN()
    results = {}
    for r in ratios:
        eng.build(xyz, offsets, k, target_ratio=r)
        results[r] = _time_ms(lambda: eng.grid.query(k, mode=1, pad_mode=1))
    best = min(results, key=results.get)
    if save:
        with _LOCK:
            t = _load()
            t[bucket(xyz.shape[0], k)] = {"target_ratio": best, "ms": results}
            tmp = TABLE_PATH + ".tmp"
            with open(tmp, "w") as f:
                json.dump(t, f, indent=1, sort_keys=True)
            os.replace(tmp, TABLE_PATH)
    return best, results
