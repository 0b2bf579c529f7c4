"""Autotuner for the grid cell size of exact queries.

The kernel's cost per query is (candidates scanned) + (merges) + (the share
of queries the first pass cannot certify, which pay a shell expansion).
Larger cells certify more queries but scan more candidates.  The cell side
is set by density feedback to give ``target_ratio * k`` candidates per
query in the 27-cell block; the best ratio depends on the cloud size, k and
how the points fill space.  ``tune`` times the exact query for a few ratios
on the actual workload, with and without two speed hints (pruning of
neighbour cells, one or two lanes per query = 64 or 32 queries per wave),
and records the choice per (log2 N, k class, density class) bucket in a
JSON table; ``lookup_*`` read it (falling back to the defaults).
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
RATIOS = (4.5, 6.5, 9.0, 13.0, 19.0, 28.0)


def _kclass(k: int) -> int:
    return 8 if k <= 8 else 16 if k <= 16 else 32 if k <= 32 else 64 if k <= 64 else 128


def bucket(n: int, k: int, dclass: int = 0) -> str:
    return f"n{max(0, int(math.log2(max(n, 1))))}_k{_kclass(k)}_d{int(dclass)}"


def _load():
    global _TABLE
    if _TABLE is None:
        try:
            with open(TABLE_PATH) as f:
                _TABLE = json.load(f)
        except (OSError, ValueError):
            _TABLE = {}
    return _TABLE


def _entry(n: int, k: int, dclass: int):
    """Table entry of the bucket; else the nearest tuned size with the same
    k class and density class; else None."""
    with _LOCK:
        t = _load()
    e = t.get(bucket(n, k, dclass))
    if e:
        return e
    suffix = f"_k{_kclass(k)}_d{int(dclass)}"
    ln = max(0, int(math.log2(max(n, 1))))
    near = [(abs(int(key[1:key.index("_")]) - ln), v) for key, v in t.items()
            if key.endswith(suffix)]
    return min(near, key=lambda x: x[0])[1] if near else None


def lookup_target_ratio(n: int, k: int, dclass: int = 0) -> float:
    """Tuned target ratio (see _entry); else the default."""
    from .knn import DEFAULT_TARGET_RATIO
    e = _entry(n, k, dclass)
    return float(e["target_ratio"]) if e else DEFAULT_TARGET_RATIO


def lookup_prune(n: int, k: int, dclass: int = 0) -> bool:
    """Tuned FK_QUERY_PRUNE choice (see _entry); else off."""
    e = _entry(n, k, dclass)
    return bool(e.get("prune", False)) if e else False


def lookup_one_lane(n: int, k: int, dclass: int = 0) -> bool:
    """Tuned FK_QUERY_ONE_LANE choice (64 instead of 32 queries per wave);
    else off."""
    e = _entry(n, k, dclass)
    return bool(e.get("one_lane", False)) if e else False


def _time_ms(fn, reps=11):
    fn()
    torch.cuda.synchronize()
    st = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    st[0].record()
    for i in range(reps):
        fn()
        st[i + 1].record()
    torch.cuda.synchronize()
    return sorted(st[i].elapsed_time(st[i + 1]) for i in range(reps))[reps // 2]


def _label(key):
    r, pr, one = key
    return f"{r}{'p' if pr else ''}{'1' if one else ''}"


def tune(xyz, offsets, k, ratios=RATIOS, save=True):
    """Time the exact query of one cloud at each target ratio, with and
    without FK_QUERY_PRUNE (k <= 16) and with one or two lanes per query
    (FK_QUERY_ONE_LANE, k <= 32); record the best under the cloud's (size, k,
    density class) bucket: the largest ratio within 3 % of the fastest.
    Returns ((ratio, prune, one_lane), {key: ms}, density class)."""
    from .knn import ExactKNN, density_class
    eng = ExactKNN()
    eng.build(xyz, offsets, k, adapt=False)
    n = xyz.shape[0]
    dclass = density_class(float(eng.grid.cloud_candidates()[0]) / max(n, 1), k)
    results = {}
    for r in ratios:
        eng.build(xyz, offsets, k, target_ratio=r)
        for pr in ((False, True) if k <= 16 else (False,)):
            for one in ((False, True) if k <= 32 else (False,)):
                results[(r, pr, one)] = _time_ms(
                    lambda: eng.grid.query(k, mode=1, pad_mode=1, prune=pr, one_lane=one))
    # within 3 % of the fastest, take the largest cells: they leave fewer
    # queries to the shell expansion, which costs most where many clouds
    # share one launch (the batched kernel)
    fastest = min(results.values())
    best = max((key for key, v in results.items() if v <= 1.03 * fastest),
               key=lambda key: (key[0], -results[key]))
    if save:
        with _LOCK:
            t = _load()
            t[bucket(n, k, dclass)] = {
                "target_ratio": best[0], "prune": best[1], "one_lane": best[2],
                "density_class": dclass,
                "ms": {_label(key): round(v, 4) for key, v in results.items()}}
            tmp = TABLE_PATH + ".tmp"
            with open(tmp, "w") as f:
                json.dump(t, f, indent=1, sort_keys=True)
            os.replace(tmp, TABLE_PATH)
    return best, results, dclass


def main(argv=None):
    """Tune the synthetic room / LiDAR workloads at a few sizes and k and
    write the table: ``python -m flashknn_exprepo_amd.autotune --out PATH``."""
    import argparse

    import numpy as np

    from . import synthetic

    global TABLE_PATH, _TABLE
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--sizes", default="100000,30000,10000")
    ap.add_argument("--ks", default="8,16,32")
    args = ap.parse_args(argv)
    TABLE_PATH, _TABLE = args.out, {}
    dev = torch.device("cuda", 0)
    bases = {"room": synthetic.room(seed=48), "lidar": synthetic.lidar(seed=49)}
    for name, base in bases.items():
        for n in (int(s) for s in args.sizes.split(",")):
            pts = synthetic.subsample_to(base, n, np.random.default_rng(n))
            x = torch.from_numpy(pts).to(dev)
            o = torch.tensor([len(pts)], dtype=torch.int32, device=dev)
            for k in (int(s) for s in args.ks.split(",")):
                best, res, dc = tune(x, o, k)
                print(f"{name} n={len(pts)} k={k} d={dc}: best ratio {best[0]}"
                      f"{' prune' if best[1] else ''}{' one-lane' if best[2] else ''}"
                      f" ({res[best]:.4f} ms; " +
                      ", ".join(f"{_label(key)}:{m:.4f}" for key, m in sorted(res.items()))
                      + ")", flush=True)


if __name__ == "__main__":
    main()
