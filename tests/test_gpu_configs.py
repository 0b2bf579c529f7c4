"""The benchmark workloads at their full size, and the hand-computed cases
of the oracle's restated rules, on the HIP path (needs an MI355X).

* LiDAR sweep (92,063 points after 5 cm voxels), adaptive neighbourhood,
  K = 16: every row bit-exact against the oracle, selected levels equal.
* 16 uneven clouds (560,035 points, sizes log-uniform in [2k, 120k], room /
  LiDAR generators seeded 50 + i, per-cloud K in {8, 16, 32}): the batched
  persistent launch, eager and replayed from its hipGraph, against the
  oracle's exact brute force (every row of clouds up to 30k points, a fixed
  20,000-row sample of the larger ones), and against the plain per-cloud
  launch on the same cells (every row).
* The adaptive level rule and back_query's truncated cells on the inputs
  whose answers are computed by hand in tests/test_oracle.py.

Bar: indices and fp32 squared distances bit-exact.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

fk = pytest.importorskip("flashknn_exprepo_amd")
from flashknn_exprepo_amd import _lib as F  # noqa: E402
from flashknn_exprepo_amd import synthetic  # noqa: E402


def _dev():
    return torch.device("cuda", 0)


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(_dev())


def _rows(tag, gi, gd, ri, rd, rows=None):
    gi = gi.cpu().numpy().astype(np.int64)
    gd = gd.cpu().numpy().astype(np.float32)
    if rows is not None:
        gi, gd = gi[rows], gd[rows]
    bad = np.where((gi != ri).any(1) | (gd.view(np.uint32) != rd.view(np.uint32)).any(1))[0]
    assert len(bad) == 0, (f"{tag}: {len(bad)} rows differ; first {bad[0]}: gpu "
                           f"{gi[bad[0]]} {gd[bad[0]]} ref {ri[bad[0]]} {rd[bad[0]]}")


# ------------------------------------------------ LiDAR sweep, adaptive

def test_lidar_sweep_adaptive_k16_full(oracle):
    lid = synthetic.lidar(seed=49)
    assert len(lid) == 92063
    grid = np.floor(lid / np.float32(0.05)).astype(np.int64)
    batch = np.zeros(len(lid), np.int64)
    ref_i, ref_d, ref_sel = oracle.adaptive_query(grid, batch, lid, 16)
    a = fk.AdaptiveNeighborhoodFlashKNN(num_nbr=16)
    idx, d = a.query(_t(grid), _t(batch), _t(lid), return_distances=True)
    _rows("adaptive", idx, d, ref_i, ref_d)
    per_level = {s["shift"]: s["query_points"] for s in a.last_stats["selected_levels"]}
    want = {int(s): int(c) for s, c in zip(*np.unique(ref_sel, return_counts=True))}
    assert per_level == want
    sel = fk.select_adaptive_levels(fk.build_adaptive_octree(_t(grid), _t(batch)), 16)
    assert np.array_equal(sel.cpu().numpy().astype(np.int64), ref_sel)


# ------------------------------------------------ 16 uneven clouds

@pytest.fixture(scope="module")
def batch16():
    clouds, ks = synthetic.batch(seed=50)
    xyz, off = synthetic.pack(clouds)
    return clouds, ks, xyz, off


def _oracle_rows(oracle, cl, k, seed):
    """(rows, idx, d): every row up to 30k points, else a fixed sample."""
    n = len(cl)
    rows = np.arange(n) if n <= 30000 else np.sort(
        np.random.default_rng(seed).choice(n, 20000, replace=False))
    ri, rd = oracle.brute_knn(cl, [n], cl[rows], [len(rows)], k)
    return rows, ri, rd


def test_batched_16_clouds_full(oracle, batch16):
    clouds, ks, xyz, off = batch16
    assert len(xyz) == 560035
    bk = fk.BatchedKNN(_t(xyz), _t(off), ks, exact=True)
    ei, ed = [t.clone() for t in bk.run()]
    bk.grid.status()
    start = 0
    for c, (cl, k) in enumerate(zip(clouds, ks)):
        bi, bd = bk.per_cloud(c)
        rows, ri, rd = _oracle_rows(oracle, cl, k, 100 + c)
        _rows(f"cloud {c} (n={len(cl)}, k={k})", bi - start, bd, ri, rd, rows)
        start += len(cl)
    # graph replay == eager, with eager work between replays
    bk.capture()
    for rep in range(3):
        bk.out_idx.fill_(-7)
        bk.out_dist.fill_(-7.0)
        torch.cuda.synchronize()
        ri, rd = bk.replay()
        torch.cuda.synchronize()
        assert torch.equal(ri, ei) and torch.equal(rd, ed), f"replay {rep}"
        _ = (torch.randn(4096, device=_dev()) * 3).sum().item()
    # the plain per-cloud launch on the same cells: every row equal
    start = 0
    for c, (cl, k) in enumerate(zip(clouds, ks)):
        e = fk.ExactKNN()
        e.build(_t(cl), _t(np.array([len(cl)], np.int32)), k, cell=bk.cell[c:c + 1],
                prune=bk.grid.prune_hint)
        oi, od = e.grid.query(k, mode=F.FK_MODE_EXACT, pad_mode=F.FK_PAD_SELF,
                              prune=e.grid.prune_hint, one_lane=e.grid.one_lane_hint)
        e.grid.status()
        bi, bd = bk.per_cloud(c)
        assert torch.equal(bi - start, oi) and torch.equal(bd, od), f"cloud {c}"
        start += len(cl)


# ------------------------------------------------ hand-computed cases

def test_adaptive_hand_levels_gpu(oracle):
    """The case of tests/test_oracle.py (levels worked out by hand from the
    reference rule): batch 0 on the x axis -- 12 points at 0, 3 at 1, one
    each at 3, 12, 15 -- and batch 1 with 8 points at 7."""
    x = [0] * 12 + [1] * 3 + [3, 12, 15] + [7] * 8
    b = np.array([0] * 18 + [1] * 8, np.int64)
    g = np.zeros((len(x), 3), np.int64)
    g[:, 0] = x
    groups = [12, 3, 1, 1, 1, 8]
    for k, want in ((1, (0, 0, 1, 2, 2, 4)), (2, (2, 2, 2, 3, 3, 4))):
        sel = fk.select_adaptive_levels(fk.build_adaptive_octree(_t(g), _t(b)), k)
        assert sel.cpu().tolist() == np.repeat(want, groups).tolist(), k
        ref_i, ref_d, _ = oracle.adaptive_query(g, b, None, k)
        a = fk.AdaptiveNeighborhoodFlashKNN(num_nbr=k)
        idx, d = a.query(_t(g), _t(b), None, return_distances=True)
        _rows(f"adaptive hand k={k}", idx, d, ref_i, ref_d)


def test_back_query_truncates_negative_cells_gpu():
    """Cells (xyz / 0.1).long(): x = 0.15 -> 1, -0.05 -> 0, -0.25 -> -2,
    -0.12 -> -1; supports are points 0 and 2 (worked out by hand in
    tests/test_oracle.py)."""
    xyz = np.zeros((4, 3), np.float32)
    xyz[:, 0] = [0.15, -0.05, -0.25, -0.12]
    sel = np.array([0, 2], np.int64)
    out = fk.FlashKNN(num_nbr=1).back_query(_t(xyz), _t(sel), 0.1, 0.05, None)
    assert out.cpu().tolist() == [0, 0, 1, 1]
    i2, d2 = fk.FlashKNN(num_nbr=2).back_query(_t(xyz), _t(sel), 0.1, 0.05, None)
    assert i2.cpu().tolist() == [[0, -1], [0, -1], [1, -1], [1, -1]]
    d2 = d2.cpu().numpy()
    assert np.isinf(d2[:, 1]).all() and d2[0, 0] == 0 and d2[2, 0] == 0
