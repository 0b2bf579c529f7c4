"""Seeded randomised parity sweep over the public query paths (needs an
MI355X): FlashKNN.query, selected_query, voxel-first subsampling,
back_query and pointops.knn_query, each bit-exact against the oracle.

Every case draws 1-3 clouds (uniform cubes of random size, room and LiDAR
subsets, quarter-grid lattices full of exact ties), a voxel size in
{2, 5, 10, 25} cm (2 cm over the LiDAR sweep takes the column layout), k,
num_down, an optional cut radius and shifted or negative cell offsets.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

fk = pytest.importorskip("flashknn_exprepo_amd")
from flashknn_exprepo_amd import synthetic  # noqa: E402

_CLOUDS = {}


def _base():
    if not _CLOUDS:
        _CLOUDS["room"] = synthetic.room(seed=48, target=30000)
        _CLOUDS["lidar"] = synthetic.lidar(seed=49)
    return _CLOUDS


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.device("cuda", 0))


def _rows_equal(tag, gi, gd, ri, rd):
    gi = gi.cpu().numpy().astype(np.int64).reshape(ri.shape)
    bad = (gi != ri).reshape(len(gi), -1).any(1)
    if gd is not None:
        g = gd.cpu().numpy().astype(np.float32).reshape(rd.shape)
        bad |= (g.view(np.uint32) != rd.view(np.uint32)).reshape(len(g), -1).any(1)
    rows = np.where(bad)[0]
    assert len(rows) == 0, f"{tag}: {len(rows)} rows differ, first {rows[:3]}"


def _cloud(rng, m):
    kind = int(rng.integers(0, 4))
    if kind == 0:
        x = rng.random((m, 3)) * rng.uniform(0.5, 10)
    elif kind == 1:
        x = synthetic.subsample_to(_base()["room"], m, rng)
    elif kind == 2:
        x = synthetic.subsample_to(_base()["lidar"], m, rng)
    else:
        x = np.round(rng.random((m, 3)) * 4) / 4
    return x.astype(np.float32)


@pytest.mark.parametrize("seed", range(24))
def test_random_case(oracle, seed):
    rng = np.random.default_rng(1000 + seed)
    B = int(rng.integers(1, 4))
    parts = [_cloud(rng, int(rng.integers(200, 8000))) for _ in range(B)]
    xyz = np.concatenate(parts)
    batch = np.concatenate([np.full(len(p), i) for i, p in enumerate(parts)]).astype(np.int64)
    vox = float(rng.choice([0.02, 0.05, 0.1, 0.25]))
    grid = np.floor(xyz / np.float32(vox)).astype(np.int64)
    if rng.random() < 0.3:
        grid = grid - grid.min(0) + rng.integers(0, 100, 3)
    nd = int(rng.integers(0, 3))
    k = int(rng.choice([1, 2, 4, 8, 16, 24, 32, 48, 64]))
    cut = math.inf if rng.random() < 0.7 else float(rng.uniform(0.05, 1.0))
    b = batch if B > 1 else None
    bt = None if b is None else _t(b)

    # FlashKNN.query (27 cells)
    ri, rd = oracle.flashknn_query(grid, b, xyz, k, nd, cut)
    knn = fk.FlashKNN(num_nbr=k, num_down=nd)
    idx = knn.query(_t(grid), bt, _t(xyz), cut_radius=cut)
    _rows_equal("query", idx, knn.last_distances, ri, rd)

    # selected_query
    qi = np.sort(rng.choice(len(xyz), int(rng.integers(1, len(xyz))), replace=False))
    ri, rd = oracle.selected_query(xyz, grid, qi, b, k, nd, cut)
    idx = knn.selected_query(_t(xyz), _t(grid), _t(qi), bt, cut_radius=cut)
    _rows_equal("selected", idx, knn.last_distances, ri, rd)

    # voxel-first subsampling + back_query on the first cloud
    x1 = parts[0]
    sel_ref = oracle.voxel_first(x1, vox)
    sel = fk.voxel_first(_t(x1), vox)
    assert np.array_equal(sel.cpu().numpy(), sel_ref), "voxel_first"
    kb = int(rng.choice([1, 3, 8]))
    qgs = vox * float(rng.choice([2.0, 4.0]))
    ri, rd = oracle.back_query(x1, sel_ref, qgs, None, kb)
    if (ri[:, 0] < 0).any():
        with pytest.raises(AssertionError):   # the reference refuses such queries
            fk.FlashKNN(num_nbr=kb).back_query(_t(x1), _t(sel_ref), qgs, vox, None)
    else:
        out = fk.FlashKNN(num_nbr=kb).back_query(_t(x1), _t(sel_ref), qgs, vox, None)
        if kb == 1:
            _rows_equal("back", out.reshape(-1, 1), None, ri[:, :1], None)
        else:
            _rows_equal("back", out[0], out[1], ri, rd)

    # pointops.knn_query (exact, sqrt distances, -1 / sqrt(1e10) padding)
    off = np.cumsum([len(p) for p in parts]).astype(np.int32)
    kk = int(rng.choice([1, 3, 16, 32, 100, 128]))
    idx, dist = fk.pointops.knn_query(kk, _t(xyz), _t(off))
    ri, rd2 = oracle.brute_knn(xyz, off, xyz, off, kk)
    rdd = np.where(ri >= 0, np.sqrt(rd2.astype(np.float32)),
                   np.float32(math.sqrt(1e10))).astype(np.float32)
    _rows_equal("pointops", idx, dist, ri, rdd)
