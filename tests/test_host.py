"""Host-side logic (CPU only): synthetic workloads, the torch helpers of the
drop-in surface, cell-size heuristics and layout conversions."""
import numpy as np
import pytest
import torch


def test_xyz2key_cpu_matches_oracle(oracle):
    from flashknn_exprepo_amd.functions import xyz2key
    rng = np.random.default_rng(0)
    g = rng.integers(-3, 1 << 16, (400, 3))
    b = rng.integers(0, 1000, 400)
    k = xyz2key(torch.from_numpy(g[:, 0]), torch.from_numpy(g[:, 1]),
                torch.from_numpy(g[:, 2]), torch.from_numpy(b))
    assert np.array_equal(k.numpy(), oracle.morton_encode(g, b))
    k8 = xyz2key(torch.tensor([255 + 256]), torch.tensor([0]), torch.tensor([0]), depth=8)
    assert k8.item() == oracle.morton_key_bitloop(255, 0, 0)


def test_cal_recall_multiset():
    from flashknn_exprepo_amd.functions import cal_recall
    gt = torch.tensor([[1, 2, 3, 4], [5, 5, 6, 7], [1, 2, 3, 4]])
    pr = torch.tensor([[4, 3, 2, 1], [5, 6, 6, 9], [9, 9, 9, 1]])
    r = cal_recall(pr, gt)
    assert torch.allclose(r, torch.tensor([1.0, 0.5, 0.25]))


def test_vanilla_knn_vs_oracle(oracle):
    from flashknn_exprepo_amd.functions import vanilla_Knn
    rng = np.random.default_rng(1)
    s = rng.random((3000, 3), dtype=np.float32)
    q = rng.random((500, 3), dtype=np.float32)
    d, i = vanilla_Knn(torch.from_numpy(q), torch.from_numpy(s), 8, 128, 1000)
    ri, rd = oracle.brute_knn(s, [3000], q, [500], 8, f64=True)
    same = (np.sort(i.numpy(), 1) == np.sort(ri, 1)).all(1).mean()
    assert same > 0.99  # matmul expansion rounding may swap near ties
    np.testing.assert_allclose(d.numpy(), np.sqrt(rd), rtol=1e-3, atol=1e-4)


def test_offsets_roundtrip():
    from flashknn_exprepo_amd import pointops
    off = torch.tensor([3, 3, 7, 12], dtype=torch.int32)
    b = pointops.offset2batch(off)
    assert b.tolist() == [0, 0, 0, 2, 2, 2, 2, 3, 3, 3, 3, 3]
    # batch2offset drops nothing for contiguous non-empty batches
    assert pointops.batch2offset(torch.tensor([0, 0, 1, 1, 1])).tolist() == [2, 5]


def test_cloud_ids_and_cell_sizes():
    from flashknn_exprepo_amd.knn import cell_sizes, cloud_ids
    off = torch.tensor([2, 5, 5, 9], dtype=torch.int32)
    assert cloud_ids(off, 9).tolist() == [0, 0, 1, 1, 1, 3, 3, 3, 3]
    rng = np.random.default_rng(2)
    xyz = torch.from_numpy(np.concatenate([
        rng.random((1000, 3), dtype=np.float32),
        rng.random((8000, 3), dtype=np.float32) * 2]))
    c = cell_sizes(xyz, torch.tensor([1000, 9000], dtype=torch.int32), 16, scale=1.0)
    # (k V / n)^(1/3): cloud 0: (16/1000)^(1/3); cloud 1: (16*8/8000)^(1/3)
    np.testing.assert_allclose(c.numpy(), [(16 / 1000) ** (1 / 3), (128 / 8000) ** (1 / 3)],
                               rtol=0.05)


def test_synthetic_uniform_and_voxel_first():
    from flashknn_exprepo_amd import synthetic
    u = synthetic.uniform(8192, 47)
    assert u.shape == (8192, 3) and u.dtype == np.float32
    assert np.array_equal(u, np.random.default_rng(47).random((8192, 3), dtype=np.float32))
    pts = np.array([[0.01, 0, 0], [0.02, 0, 0], [0.5, 0, 0], [0.03, 0, 0]], np.float32)
    assert synthetic.voxel_first(pts, 0.1).tolist() == [0, 2]


@pytest.mark.slow
def test_synthetic_room_size():
    from flashknn_exprepo_amd import synthetic
    r = synthetic.room(seed=48, target=20000)
    assert abs(len(r) - 20000) <= 0.05 * 20000
    g = np.floor(r / 0.04).astype(np.int64)
    assert len(np.unique(g, axis=0)) == len(r)  # one point per 4 cm voxel


def test_pack_offsets():
    from flashknn_exprepo_amd import synthetic
    xyz, off = synthetic.pack([np.zeros((3, 3)), np.ones((5, 3))])
    assert off.tolist() == [3, 8] and xyz.shape == (8, 3)


def test_autotune_lookup_and_density_class(tmp_path, monkeypatch):
    import json
    from flashknn_exprepo_amd import autotune
    from flashknn_exprepo_amd.knn import DEFAULT_TARGET_RATIO, density_class
    assert density_class(16 * 9, 16) == 3 and density_class(3, 16) == 0
    table = {autotune.bucket(100000, 16, 3): {"target_ratio": 13.0},
             autotune.bucket(10000, 16, 3): {"target_ratio": 6.5}}
    p = tmp_path / "t.json"
    p.write_text(json.dumps(table))
    monkeypatch.setattr(autotune, "TABLE_PATH", str(p))
    monkeypatch.setattr(autotune, "_TABLE", None)
    assert autotune.lookup_target_ratio(100000, 16, 3) == 13.0
    assert autotune.lookup_target_ratio(12000, 16, 3) == 6.5      # nearest size
    assert autotune.lookup_target_ratio(100000, 32, 3) == DEFAULT_TARGET_RATIO
    assert autotune.lookup_prune(100000, 16, 3) is False             # absent: off
    table[autotune.bucket(100000, 16, 5)] = {"target_ratio": 19.0, "prune": True}
    p.write_text(json.dumps(table))
    monkeypatch.setattr(autotune, "_TABLE", None)
    assert autotune.lookup_prune(90000, 16, 5) is True
    assert autotune.lookup_target_ratio(90000, 16, 5) == 19.0
    # a far size of the same density class loses to a near size one class off
    table = {autotune.bucket(1500, 16, 5): {"target_ratio": 28.0},
             autotune.bucket(90000, 16, 4): {"target_ratio": 9.0, "prune": True}}
    p.write_text(json.dumps(table))
    monkeypatch.setattr(autotune, "_TABLE", None)
    assert autotune.lookup_target_ratio(250000, 16, 5) == 9.0
    assert autotune.lookup_prune(250000, 16, 5) is True
    assert autotune.lookup_target_ratio(2000, 16, 5) == 28.0


def test_quotient32_correctly_rounded():
    """fl32(x / fl32(size)) through float64 (double rounding is innocuous
    for a quotient of fp32 values) equals numpy's fp32 division bitwise."""
    from flashknn_exprepo_amd.functions.FlashKnnWrapper import quotient32
    rng = np.random.default_rng(3)
    x = np.concatenate([np.array([5.0399995, -7.5599995, 0.0, -0.0, 1e-40, 3e38], np.float32),
                        (rng.standard_normal(200000) * 10).astype(np.float32)])
    for size in (0.06, 0.04, 0.24, 1e-3, 7.0):
        got = quotient32(torch.from_numpy(x), size).numpy()
        assert np.array_equal(got.view(np.uint32), (x / np.float32(size)).view(np.uint32)), size


@pytest.mark.parametrize("lo,hi", [(32, 128), (16, 16), (100, 40)])
def test_walk_levels_equals_the_level_loop(lo, hi):
    """The vectorised coarse-to-fine walk of the adaptive neighbourhood equals
    the level-by-level loop it replaces (including lo > hi)."""
    from flashknn_exprepo_amd.functions.adaptive_octree import walk_levels
    rng = np.random.default_rng(lo * 7 + hi)
    for L in (1, 2, 5, 9):
        shifts = torch.arange(11, 11 - L, -1)
        # counts shrink towards finer levels, with noise, hitting both bounds
        base = rng.integers(0, 400, 5000)
        c = np.stack([base >> (11 - int(s)) // 2 for s in shifts]) + rng.integers(0, 60, (L, 5000))
        c = torch.from_numpy(np.sort(c, axis=0)[::-1].copy()).to(torch.int32)
        c[:, :50] = torch.from_numpy(rng.integers(0, 300, (L, 50))).to(torch.int32)
        sel = torch.full((5000,), int(shifts[0]), dtype=torch.int16)
        active = c[0] > hi
        for i in range(1, L):
            use = active & (c[i] >= lo)
            sel = torch.where(use, torch.full_like(sel, int(shifts[i])), sel)
            active = use & (c[i] > hi)
        s2, a2 = walk_levels(c, shifts, lo, hi)
        assert torch.equal(s2, sel) and torch.equal(a2, active), L


@pytest.mark.parametrize("n,B,contiguous", [(1, 1, True), (3000, 1, True), (5000, 4, True),
                                            (5000, 4, False), (70000, 9, True), (2500, 6, True)])
def test_batch_min_max_cpu(n, B, contiguous):
    """Per-batch minimum / maximum (torch ops, any device) against numpy,
    including chunk padding, interleaved batches and batches without points."""
    from flashknn_exprepo_amd.functions.adaptive_octree import batch_max, batch_min
    rng = np.random.default_rng(n * 3 + B)
    b = rng.integers(0, B - 1 if B == 6 else B, n)          # B == 6: last batch empty
    if contiguous:
        b.sort()
    v = rng.integers(-10**12, 10**12, (n, 3))
    lo = np.full((B, 3), np.iinfo(np.int64).max, np.int64)
    hi = np.full((B, 3), np.iinfo(np.int64).min, np.int64)
    np.minimum.at(lo, b, v)
    np.maximum.at(hi, b, v)
    tv, tb = torch.from_numpy(v.astype(np.int64)), torch.from_numpy(b.astype(np.int64))
    assert np.array_equal(batch_min(tv, tb, B).numpy(), lo)
    assert np.array_equal(batch_max(tv, tb, B).numpy(), hi)
