"""The adaptive-neighbourhood query on the device (fk_adaptive_select +
fk_adaptive_query, k <= 64; the per-level grid path for k > 64), against the
oracle's restatement of FlashKNN/functions/adaptive_octree.py (needs an
MI355X).  Bar: indices and fp32 squared distances bit-exact, selected
levels equal, per-level statistics equal to the oracle's counts.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

fk = pytest.importorskip("flashknn_exprepo_amd")
from flashknn_exprepo_amd import synthetic  # noqa: E402


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.device("cuda", 0))


def _same(idx, d, ri, rd):
    gi = idx.cpu().numpy()
    gd = d.cpu().numpy()
    bad = np.where((gi != ri).any(1) | (gd.view(np.uint32) != rd.view(np.uint32)).any(1))[0]
    assert len(bad) == 0, (f"{len(bad)} rows differ; first {bad[0]}: {gi[bad[0]]} {gd[bad[0]]}"
                           f" vs {ri[bad[0]]} {rd[bad[0]]}")


def _lidar(n, seed=49):
    pts = synthetic.lidar_points(n, seed)
    return pts, np.floor(pts / np.float32(0.05)).astype(np.int64)


@pytest.mark.parametrize("k", [4, 16, 32, 64])
def test_device_path_levels_and_rows(oracle, k):
    pts, grid = _lidar(12000)
    batch = np.zeros(len(pts), np.int64)
    ref_i, ref_d, ref_sel = oracle.adaptive_query(grid, batch, pts, k)
    a = fk.AdaptiveNeighborhoodFlashKNN(num_nbr=k)
    print(f"adaptive device path k={k}", flush=True)
    idx, d = a.query(_t(grid), _t(batch), _t(pts), return_distances=True)
    _same(idx, d, ref_i, ref_d)
    assert np.array_equal(a.last_levels.cpu().numpy(), ref_sel)
    # per-level statistics: the selected points' 27-node counts
    g = grid - grid.min(0)
    counts = {s: oracle.point_candidates(g >> s, batch) for s in np.unique(ref_sel)}
    for st in a.last_stats["selected_levels"]:
        m = ref_sel == st["shift"]
        c = counts[st["shift"]][m]
        assert st["query_points"] == m.sum()
        assert (st["candidate_min"], st["candidate_max"]) == (c.min(), c.max())
        assert abs(st["candidate_mean"] - c.mean()) < 1e-6 * c.mean()


def test_device_path_max_shift_cut_and_batches(oracle):
    pts, grid = _lidar(9000)
    batch = (np.arange(len(pts)) * 3 // len(pts)).astype(np.int64)
    grid = grid + 5000 * batch[:, None] - 10**6
    for max_shift, cut in ((9, float("inf")), (12, 0.6), (None, 0.3)):
        ref_i, ref_d, ref_sel = oracle.adaptive_query(grid, batch, pts, 16,
                                                      max_shift=max_shift, cut_radius=cut)
        a = fk.AdaptiveNeighborhoodFlashKNN(num_nbr=16)
        print(f"adaptive max_shift={max_shift} cut={cut}", flush=True)
        idx, d = a.query(_t(grid), _t(batch), _t(pts), max_shift=max_shift, cut_radius=cut,
                         return_distances=True)
        _same(idx, d, ref_i, ref_d)
        assert np.array_equal(a.last_levels.cpu().numpy(), ref_sel)


def test_levels_path_above_64(oracle):
    pts, grid = _lidar(5000)
    batch = np.zeros(len(pts), np.int64)
    ref_i, ref_d, _ = oracle.adaptive_query(grid, batch, pts, 80)
    a = fk.AdaptiveNeighborhoodFlashKNN(num_nbr=80)
    idx, d = a.query(_t(grid), _t(batch), _t(pts), return_distances=True)
    _same(idx, d, ref_i, ref_d)


def test_device_path_errors():
    rng = np.random.default_rng(4)
    grid = rng.integers(0, 100, (500, 3)).astype(np.int64)
    a = fk.AdaptiveNeighborhoodFlashKNN(num_nbr=8)
    bad_b = np.zeros(500, np.int64)
    bad_b[7] = 40000
    with pytest.raises(ValueError, match="batch indices"):
        a.query(_t(grid), _t(bad_b), None)
    wide = grid.copy()
    wide[3, 1] = 70000
    with pytest.raises(ValueError, match="span"):
        a.query(_t(wide), _t(np.zeros(500, np.int64)), None)
    with pytest.raises(ValueError, match="fewer than"):
        fk.AdaptiveNeighborhoodFlashKNN(num_nbr=16).query(_t(grid[:10]),
                                                           _t(np.zeros(10, np.int64)), None)
    with pytest.raises(ValueError, match="max_shift"):
        a.query(_t(grid), _t(np.zeros(500, np.int64)), None, max_shift=17)
    # a good call on the same object still works after the errors
    idx = a.query(_t(grid), _t(np.zeros(500, np.int64)), None)
    assert idx.shape == (500, 8)
