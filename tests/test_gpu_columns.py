"""Column layout of the grid (needs an MI355X).

A build whose dense cell boxes exceed 2^30 cells lays the cloud out by (x, y)
columns, with the points of a column sorted by z (FK_BUILD_COLUMNS forces it
at any size).  Candidate sets, and so results, are those of the dense layout:
every query path is checked bit-exact against the oracle, first with the
layout forced on small clouds, then with the fallback taken by itself (2 cm
cells over a 100 m LiDAR sweep: ~7e9 dense cells, 2.5e7 columns).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

fk = pytest.importorskip("flashknn_exprepo_amd")
from flashknn_exprepo_amd import _lib as F  # noqa: E402
from flashknn_exprepo_amd import synthetic  # noqa: E402
from flashknn_exprepo_amd.grid import Grid  # noqa: E402


def _dev():
    return torch.device("cuda", 0)


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(_dev())


def _same(gpu_idx, gpu_d, ref_idx, ref_d):
    gi = gpu_idx.cpu().numpy().astype(np.int64)
    gd = gpu_d.cpu().numpy().astype(np.float32)
    bad = np.where((gi != ref_idx).any(1) | (gd.view(np.uint32) != ref_d.view(np.uint32)).any(1))[0]
    assert len(bad) == 0, (
        f"{len(bad)} rows differ; first row {bad[0]}: gpu {gi[bad[0]]} {gd[bad[0]]} "
        f"ref {ref_idx[bad[0]]} {ref_d[bad[0]]}")


def _columns(cells, batch):
    """Table entries of the column layout: sum over clouds of the x-y box."""
    tot = 0
    for b in np.unique(batch):
        c = cells[batch == b]
        ext = c.max(0) - c.min(0) + 1
        tot += int(ext[0]) * int(ext[1])
    return tot


def _room_two_clouds():
    xyz = synthetic.room(seed=48, target=20000)
    batch = (np.random.default_rng(62).random(len(xyz)) < 0.4).astype(np.int64)
    return xyz, batch


# ------------------------------------------------------ forced layout

@pytest.mark.parametrize("num_down", [0, 1])
@pytest.mark.parametrize("k", [1, 16, 32, 100])
def test_columns_cells_query(oracle, k, num_down):
    xyz, batch = _room_two_clouds()
    grid = np.floor(xyz / np.float32(0.04)).astype(np.int64)
    ref_i, ref_d = oracle.flashknn_query(grid, batch, xyz, k, num_down)
    g = Grid().build_cells(_t(grid), _t(batch), _t(xyz), num_down=num_down, columns=True)
    assert g.info()["cells"] == _columns(grid >> num_down, batch)
    idx, d = g.query(k)
    _same(idx, d, ref_i, ref_d)
    # the dense layout of the same input agrees
    g2 = Grid().build_cells(_t(grid), _t(batch), _t(xyz), num_down=num_down)
    assert g2.info()["cells"] > g.info()["cells"]
    i2, d2 = g2.query(k)
    assert torch.equal(i2, idx) and torch.equal(d2, d)


def test_columns_selected_and_cut(oracle):
    xyz, _ = _room_two_clouds()
    grid = np.floor(xyz / np.float32(0.04)).astype(np.int64)
    qi = np.random.default_rng(63).choice(len(xyz), 3000, replace=False)
    ref_i, ref_d = oracle.selected_query(xyz, grid, qi, None, 16, 1, 0.1)
    g = Grid().build_cells(_t(grid), None, _t(xyz), num_down=1, columns=True)
    g.attach_queries(_t(xyz[qi]), qcell=_t(grid[qi]), qpad=_t(qi.astype(np.int32)))
    idx, d = g.query(16, cut_r2=0.1 ** 2, pad_mode=F.FK_PAD_PER_QUERY, attached=True)
    _same(idx, d, ref_i, ref_d)


@pytest.mark.parametrize("k", [1, 16, 64, 100])
def test_columns_exact(oracle, k):
    rng = np.random.default_rng(64)
    a = synthetic.subsample_to(synthetic.room(seed=48, target=30000), 12000, rng)
    b = synthetic.subsample_to(synthetic.lidar(seed=49), 15000, rng)
    xyz, off = synthetic.pack([a, b])
    ref_i, ref_d = oracle.brute_knn(xyz, off, xyz, off, k)
    inv = torch.tensor([1 / 0.05, 1 / 0.2], dtype=torch.float32, device=_dev())
    g = Grid().build_xyz(_t(xyz), _t(off), inv, columns=True)
    for prune in ((False, True) if k <= 16 else (False,)):
        idx, d = g.query(k, mode=F.FK_MODE_EXACT, pad_mode=F.FK_PAD_MINUS1, prune=prune)
        _same(idx, d, ref_i, ref_d)


def test_columns_exact_other_queries(oracle):
    rng = np.random.default_rng(65)
    s = rng.random((3000, 3), dtype=np.float32)
    q = (rng.random((900, 3), dtype=np.float32) * 1.4 - 0.2).astype(np.float32)
    soff = np.array([5, 1500, 3000], np.int32)
    qoff = np.array([300, 600, 900], np.int32)
    ref_i, ref_d = oracle.brute_knn(s, soff, q, qoff, 8)
    inv = torch.full((3,), 1 / 0.07, dtype=torch.float32, device=_dev())
    g = Grid().build_xyz(_t(s), _t(soff), inv, columns=True)
    g.attach_queries(_t(q), qoffsets=_t(qoff))
    idx, d = g.query(8, mode=F.FK_MODE_EXACT, pad_mode=F.FK_PAD_MINUS1, attached=True)
    _same(idx, d, ref_i, ref_d)


@pytest.mark.parametrize("exact", [True, False])
def test_columns_batched(exact):
    rng = np.random.default_rng(66)
    clouds = [synthetic.subsample_to(synthetic.room(seed=48, target=30000), 9000, rng),
              synthetic.subsample_to(synthetic.lidar(seed=49), 14000, rng),
              rng.random((2500, 3), dtype=np.float32)]
    xyz, off = synthetic.pack(clouds)
    bk = fk.BatchedKNN(_t(xyz), _t(off), [8, 16, 32], exact=exact)
    ei, ed = [t.clone() for t in bk.run()]
    bk.grid.build_xyz(bk.xyz, bk.offsets, (1.0 / bk.cell.double()).float(), columns=True)
    ci, cd = bk.run()
    assert torch.equal(ci, ei) and torch.equal(cd, ed)


def test_columns_refused_without_sync():
    g = Grid()
    g.reserve(100, 1000, 1)
    x = torch.zeros((4, 3), dtype=torch.int64, device=_dev())
    with pytest.raises(ValueError):      # FK_ERR_INVALID_ARG
        g.build_cells(x, no_sync=True, columns=True)


# ------------------------------------------- fallback beyond 2^30 cells

def _lidar_2cm(n=20000, seed=67):
    rng = np.random.default_rng(seed)
    pts = synthetic.subsample_to(synthetic.lidar(seed=49), n, rng)
    grid = np.floor(pts / np.float32(0.02)).astype(np.int64)
    ext = grid.max(0) - grid.min(0) + 1
    assert np.prod(ext.astype(np.float64)) > 2.0 ** 30
    return pts, grid


@pytest.mark.parametrize("k", [1, 16, 64])
def test_fallback_query(oracle, k):
    pts, grid = _lidar_2cm()
    ref_i, ref_d = oracle.flashknn_query(grid, None, pts, k, 0)
    knn = fk.FlashKNN(num_nbr=k, num_down=0)
    idx = knn.query(_t(grid), None, _t(pts))
    _same(idx, knn.last_distances, ref_i, ref_d)


def test_fallback_selected_and_back(oracle):
    pts, grid = _lidar_2cm()
    qi = np.random.default_rng(68).choice(len(pts), 4000, replace=False)
    ref_i, ref_d = oracle.selected_query(pts, grid, qi, None, 16, 0)
    knn = fk.FlashKNN(num_nbr=16, num_down=0)
    idx = knn.selected_query(_t(pts), _t(grid), _t(qi), None)
    _same(idx, knn.last_distances, ref_i, ref_d)
    sel = synthetic.voxel_first(pts, 0.02)
    ref_i, ref_d = oracle.back_query(pts, sel, 0.02, None, 4)
    out = fk.FlashKNN(num_nbr=4).back_query(_t(pts), _t(sel), 0.02, 0.02, None)
    _same(out[0], out[1], ref_i, ref_d)


def test_fallback_exact_explicit_cell(oracle):
    pts, _ = _lidar_2cm(15000, 69)
    off = np.array([len(pts)], np.int32)
    ref_i, ref_d = oracle.brute_knn(pts, off, pts, off, 16)
    idx, d = fk.knn(_t(pts), _t(off), 16, cell=torch.tensor([0.02]))
    _same(idx, d, ref_i, ref_d)


def test_fallback_adaptive(oracle):
    pts, grid = _lidar_2cm(12000, 70)
    batch = np.zeros(len(pts), np.int64)
    ref_i, ref_d, _ = oracle.adaptive_query(grid, batch, pts, 16)
    a = fk.AdaptiveNeighborhoodFlashKNN(num_nbr=16)
    idx, d = a.query(_t(grid), _t(batch), _t(pts), return_distances=True)
    _same(idx, d, ref_i, ref_d)


def test_fallback_limit():
    """Columns are capped like cells: an x-y extent beyond 2^30 columns raises."""
    knn = fk.FlashKNN(num_nbr=4, num_down=0)
    g = torch.tensor([[0, 0, 0], [40000, 40000, 7]], dtype=torch.int64, device=_dev())
    with pytest.raises(fk.FlashKNNError):
        knn.query(g, None, None)
