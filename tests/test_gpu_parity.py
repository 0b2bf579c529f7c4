"""Parity of the HIP path against the CPU oracle (needs an MI355X).

Bar: neighbour indices and fp32 squared distances bit-exact; sets compared
row by row where the reference's tie rule differs (the build breaks exact
ties by the smaller index, which the oracle does too, so rows are compared
exactly).
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

fk = pytest.importorskip("flashknn_exprepo_amd")


def _dev():
    return torch.device("cuda", 0)


def _t(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(_dev())
    return t if dtype is None else t.to(dtype)


def _same(gpu_idx, gpu_d, ref_idx, ref_d):
    gi = gpu_idx.cpu().numpy().astype(np.int64)
    gd = gpu_d.cpu().numpy().astype(np.float32)
    bad = np.where((gi != ref_idx).any(1) | (gd.view(np.uint32) != ref_d.view(np.uint32)).any(1))[0]
    assert len(bad) == 0, (
        f"{len(bad)} rows differ; first row {bad[0]}: gpu {gi[bad[0]]} {gd[bad[0]]} "
        f"ref {ref_idx[bad[0]]} {ref_d[bad[0]]}")


# ------------------------------------------------ FlashKNN.query (27 cells)

@pytest.mark.parametrize("k", [1, 2, 8, 16, 24, 32, 48, 64, 96, 128])
def test_query_uniform_cells(oracle, k):
    xyz = np.random.default_rng(47).random((8192, 3), dtype=np.float32)
    grid = np.floor(xyz / np.float32(0.05)).astype(np.int64)
    ref_i, ref_d = oracle.flashknn_query(grid, None, xyz, k, 0)
    knn = fk.FlashKNN(num_nbr=k, num_down=0)
    idx = knn.query(_t(grid), None, _t(xyz))
    assert idx.dtype == torch.int64 and idx.shape == (8192, k)
    _same(idx, knn.last_distances, ref_i, ref_d)


@pytest.mark.parametrize("num_down", [0, 1, 2])
def test_query_room_num_down(oracle, num_down):
    from flashknn_exprepo_amd import synthetic
    xyz = synthetic.room(seed=48, target=20000)
    grid = np.floor(xyz / np.float32(0.04)).astype(np.int64)
    ref_i, ref_d = oracle.flashknn_query(grid, None, xyz, 16, num_down)
    knn = fk.FlashKNN(num_nbr=16, num_down=num_down)
    idx = knn.query(_t(grid), None, _t(xyz))
    _same(idx, knn.last_distances, ref_i, ref_d)


def test_query_batches_and_negative_cells(oracle):
    rng = np.random.default_rng(3)
    xyz = (rng.random((6000, 3), dtype=np.float32) * 4 - 2).astype(np.float32)
    grid = np.floor(xyz / np.float32(0.1)).astype(np.int64)
    batch = rng.integers(0, 5, 6000)
    ref_i, ref_d = oracle.flashknn_query(grid, batch, xyz, 12, 1)
    knn = fk.FlashKNN(num_nbr=12, num_down=1)
    idx = knn.query(_t(grid), _t(batch), _t(xyz))
    _same(idx, knn.last_distances, ref_i, ref_d)


def test_query_duplicates_ties_and_cut(oracle):
    rng = np.random.default_rng(4)
    base = np.round(rng.random((500, 3)) * 8).astype(np.float32) / 8
    xyz = np.concatenate([base, base[:200], base[:50]]).astype(np.float32)
    grid = np.floor(xyz / np.float32(0.125)).astype(np.int64)
    for cut in (math.inf, 0.2):
        ref_i, ref_d = oracle.flashknn_query(grid, None, xyz, 16, 1, cut)
        knn = fk.FlashKNN(num_nbr=16, num_down=1)
        idx = knn.query(_t(grid), None, _t(xyz), cut_radius=cut)
        _same(idx, knn.last_distances, ref_i, ref_d)


def test_query_xyz_none_uses_grid(oracle):
    rng = np.random.default_rng(5)
    grid = rng.integers(0, 40, (3000, 3))
    ref_i, ref_d = oracle.flashknn_query(grid, None, None, 8, 2)
    knn = fk.FlashKNN(num_nbr=8, num_down=2)
    idx = knn.query(_t(grid), None, None)
    _same(idx, knn.last_distances, ref_i, ref_d)


def test_query_tiny_and_empty(oracle):
    xyz = np.array([[0.1, 0.1, 0.1]], np.float32)
    grid = np.zeros((1, 3), np.int64)
    knn = fk.FlashKNN(num_nbr=4, num_down=0)
    idx = knn.query(_t(grid), None, _t(xyz))
    assert idx.cpu().tolist() == [[0, 0, 0, 0]]
    d = knn.last_distances.cpu().numpy()
    assert d[0, 0] == 0 and np.isinf(d[0, 1:]).all()
    idx0 = knn.query(_t(np.zeros((0, 3), np.int64)), None, _t(np.zeros((0, 3), np.float32)))
    assert idx0.shape == (0, 4)


def test_query_large_sparse_box(oracle):
    """Clusters spread over 60 m x 60 m x 10 m at 5 cm cells: a dense box of
    ~3.3e8 cells (above 2^28, below the 2^30 cap) builds and matches; boxes
    beyond the cap (one cloud, or the total over clouds) take the column
    layout and match too."""
    rng = np.random.default_rng(61)
    centres = rng.random((6, 3)) * np.array([60.0, 60.0, 10.0])
    xyz = np.concatenate([c + rng.normal(0, 0.15, (800, 3)) for c in centres] +
                         [np.array([[0.01, 0.01, 0.01], [59.99, 59.99, 9.99]])])
    xyz = xyz.astype(np.float32)
    grid = np.floor(xyz / np.float32(0.05)).astype(np.int64)
    span = grid.max(0) - grid.min(0) + 1
    assert 1 << 28 < np.prod(span) < 1 << 30
    for k in (1, 16):
        ref_i, ref_d = oracle.flashknn_query(grid, None, xyz, k, 0)
        knn = fk.FlashKNN(num_nbr=k, num_down=0)
        idx = knn.query(_t(grid), None, _t(xyz))
        _same(idx, knn.last_distances, ref_i, ref_d)
    knn = fk.FlashKNN(num_nbr=8, num_down=0)
    one = np.concatenate([grid, [[1100, 1100, 1100]]])
    half = np.concatenate([grid, [[0, 0, 600]]])
    two = np.concatenate([half, half + 3])
    bt = np.repeat([0, 1], len(half))
    box = lambda c: float(np.prod((c.max(0) - c.min(0) + 1).astype(np.float64)))
    assert box(one) > 2.0 ** 30                                    # one cloud
    assert box(half) < 2.0 ** 30 < 2 * box(half)                   # the total
    for g, b in ((one, None), (two, bt)):
        ref_i, ref_d = oracle.flashknn_query(g, b, None, 8, 0)
        idx = knn.query(_t(g), None if b is None else _t(b), None)
        _same(idx, knn.last_distances, ref_i, ref_d)


def test_query_mode_validation():
    knn = fk.FlashKNN(num_nbr=4, num_down=0)
    g = _t(np.zeros((4, 3), np.int64))
    with pytest.raises(ValueError):
        knn.query(g, None, None, candidate_mode="bogus")
    with pytest.raises(ValueError):
        knn.query(g, None, None, thread_group_size=12)
    with pytest.raises(ValueError):
        fk.FlashKNN(num_nbr=129, num_down=0).query(g, None, None)
    with pytest.raises(ValueError):
        knn.query(g, _t(np.array([0, 1, -1, 0])), None)


def test_query_modes_identical(oracle):
    xyz = np.random.default_rng(6).random((4000, 3), dtype=np.float32)
    grid = np.floor(xyz / np.float32(0.04)).astype(np.int64)
    knn = fk.FlashKNN(num_nbr=16, num_down=1)
    a = knn.query(_t(grid), None, _t(xyz))
    for kw in ({"memory_mode": "GM", "sorting_mode": "SS"}, {"thread_group_size": 8},
               {"candidate_mode": "shared"}, {"enable_skip": False}):
        b = knn.query(_t(grid), None, _t(xyz), **kw)
        assert torch.equal(a, b)


def test_traverse_info(oracle):
    xyz = np.random.default_rng(7).random((3000, 3), dtype=np.float32)
    grid = np.floor(xyz / np.float32(0.1)).astype(np.int64)
    info = [None]
    fk.FlashKNN(num_nbr=8, num_down=0).query(_t(grid), None, _t(xyz), traverse_info=info)
    # brute count of the 27-cell candidate sets
    from collections import Counter
    cnt = Counter(map(tuple, grid))
    tot = 0
    for g in grid:
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    tot += cnt.get((g[0] + dx, g[1] + dy, g[2] + dz), 0)
    assert abs(info[0] - tot / len(grid)) < 1e-9


# ------------------------------------------- selected_query / back_query

def test_selected_query(oracle):
    rng = np.random.default_rng(8)
    xyz = rng.random((5000, 3), dtype=np.float32)
    grid = np.floor(xyz / np.float32(0.05)).astype(np.int64)
    qi = rng.choice(5000, 700, replace=False)
    ref_i, ref_d = oracle.selected_query(xyz, grid, qi, None, 16, 1)
    knn = fk.FlashKNN(num_nbr=16, num_down=1)
    idx = knn.selected_query(_t(xyz), _t(grid), _t(qi), None)
    _same(idx, knn.last_distances, ref_i, ref_d)


@pytest.mark.parametrize("k", [1, 4])
def test_back_query(oracle, k):
    from flashknn_exprepo_amd import synthetic
    xyz = synthetic.room(seed=48, target=20000)
    sel = synthetic.voxel_first(xyz, 0.08)
    ref_i, ref_d = oracle.back_query(xyz, sel, 0.16, None, k)
    out = fk.FlashKNN(num_nbr=k).back_query(_t(xyz), _t(sel), 0.16, 0.08, None)
    if k == 1:
        assert torch.equal(out.cpu(), torch.from_numpy(ref_i[:, 0]))
    else:
        _same(out[0], out[1], ref_i, ref_d)


# ------------------------------------------------------- exact kNN

@pytest.mark.parametrize("k", [1, 8, 16, 32, 64, 65, 100])
def test_exact_uniform_vs_brute(oracle, k):
    xyz = np.random.default_rng(47).random((8192, 3), dtype=np.float32)
    ref_i, ref_d = oracle.brute_knn(xyz, [8192], xyz, [8192], k)
    idx, d = fk.knn(_t(xyz), _t(np.array([8192], np.int32)), k)
    _same(idx, d, ref_i, ref_d)


@pytest.mark.parametrize("cell", [0.03, 0.15, 0.8, None])
def test_exact_room_lidar_any_cell(oracle, cell):
    from flashknn_exprepo_amd import synthetic
    rng = np.random.default_rng(9)
    a = synthetic.subsample_to(synthetic.room(seed=48, target=30000), 12000, rng)
    b = synthetic.subsample_to(synthetic.lidar(seed=49), 15000, rng)
    xyz, off = synthetic.pack([a, b])
    ref_i, ref_d = oracle.brute_knn(xyz, off, xyz, off, 16)
    eng = fk.ExactKNN()
    c = None if cell is None else torch.tensor([cell, cell * 2], device=_dev())
    idx, d = eng(_t(xyz), _t(off), 16, cell=c, stats=True)
    _same(idx, d, ref_i, ref_d)
    st = eng.last_stats
    assert st["certified_first_pass"] + st["expanded"] == len(xyz)


def test_exact_other_queries_and_padding(oracle):
    rng = np.random.default_rng(10)
    s = rng.random((3000, 3), dtype=np.float32)
    q = (rng.random((900, 3), dtype=np.float32) * 1.4 - 0.2).astype(np.float32)
    soff = np.array([5, 1500, 3000], np.int32)      # a 5-point cloud: padding
    qoff = np.array([300, 600, 900], np.int32)
    ref_i, ref_d = oracle.brute_knn(s, soff, q, qoff, 8)
    idx, d = fk.knn(_t(s), _t(soff), 8, _t(q), _t(qoff))
    _same(idx, d, ref_i, ref_d)
    assert (idx[:300, 5:] == -1).all()


def test_exact_engine_grown_cells_then_no_sync_attach(oracle):
    """A synchronising build (explicit fine cell) grows the cell tables past
    the engine's reserve; the next no-sync build + attach still fits."""
    rng = np.random.default_rng(71)
    s = rng.random((3000, 3), dtype=np.float32)
    q = rng.random((700, 3), dtype=np.float32)
    off, qoff = np.array([3000], np.int32), np.array([700], np.int32)
    eng = fk.ExactKNN()
    eng(_t(s), _t(off), 8)                                   # reserves, no-sync
    idx, d = eng(_t(s), _t(off), 8, cell=torch.tensor([0.004]))   # 1.6e7 cells, sync
    ref_i, ref_d = oracle.brute_knn(s, off, s, off, 8)
    _same(idx, d, ref_i, ref_d)
    idx, d = eng(_t(s), _t(off), 8, _t(q), _t(qoff))         # no-sync build + attach
    ref_i, ref_d = oracle.brute_knn(s, off, q, qoff, 8)
    _same(idx, d, ref_i, ref_d)


def test_reserved_cloud_count_is_exact(oracle):
    """A no-sync build may use up to the reserved number of clouds (their box
    buffers included) and is refused beyond it; a handle reserved for many
    clouds builds fewer (16 box copies each) correctly."""
    from flashknn_exprepo_amd.grid import Grid
    rng = np.random.default_rng(72)
    n = 6000
    grid = rng.integers(0, 4, (n, 3))
    g = Grid()
    g.reserve(n, 1 << 16, 10, n)
    for B in (1, 10):
        batch = rng.integers(0, B, n)
        g.build_cells(_t(grid), _t(batch), None, num_clouds=B, no_sync=True)
        g.status()
        idx, d = g.query(8)
        ref_i, ref_d = oracle.flashknn_query(grid, batch, None, 8, 0)
        _same(idx, d, ref_i, ref_d)
    with pytest.raises(fk.FlashKNNError):
        g.build_cells(_t(grid), _t(rng.integers(0, 50, n)), None, num_clouds=50,
                      no_sync=True)
    g2 = Grid()
    g2.reserve(n, 1 << 18, 2000, n)
    batch = rng.integers(0, 1000, n)
    g2.build_cells(_t(grid), _t(batch), None, num_clouds=1000, no_sync=True)
    g2.status()
    idx, d = g2.query(8)
    ref_i, ref_d = oracle.flashknn_query(grid, batch, None, 8, 0)
    _same(idx, d, ref_i, ref_d)


def test_pointops_knn_query(oracle):
    rng = np.random.default_rng(11)
    s = rng.random((4000, 3), dtype=np.float32)
    off = np.array([1000, 4000], np.int32)
    ref_i, ref_d = oracle.brute_knn(s, off, s, off, 16)
    idx, dist = fk.pointops.knn_query(16, _t(s), _t(off))
    assert idx.dtype == torch.int32 and dist.dtype == torch.float32
    assert np.array_equal(idx.cpu().numpy(), ref_i)
    np.testing.assert_array_equal(dist.cpu().numpy(), np.sqrt(ref_d))


@pytest.mark.parametrize("k", [65, 100, 128])
def test_pointops_knn_query_wide(oracle, k):
    """nsample above 64 (one wave per query) up to the reference's 128; the
    90-point segment pads with (-1, sqrt(1e10)); attached queries too."""
    rng = np.random.default_rng(17)
    s = rng.random((5000, 3), dtype=np.float32)
    off = np.array([90, 2000, 5000], np.int32)
    q = rng.random((700, 3), dtype=np.float32)
    qoff = np.array([100, 400, 700], np.int32)
    for new_xyz, new_off in ((None, None), (q, qoff)):
        qq, qo = (s, off) if new_xyz is None else (new_xyz, new_off)
        ref_i, ref_d = oracle.brute_knn(s, off, qq, qo, k)
        if new_xyz is None:
            idx, dist = fk.pointops.knn_query(k, _t(s), _t(off))
        else:
            idx, dist = fk.pointops.knn_query(k, _t(s), _t(off), _t(q), _t(qoff))
        i_np, d_np = idx.cpu().numpy(), dist.cpu().numpy()
        assert np.array_equal(i_np, ref_i)
        pad = ref_i < 0
        assert not pad[:qo[0], :90].any() and pad[:qo[0], 90:].all()
        assert not pad[qo[0]:].any()
        np.testing.assert_array_equal(d_np[~pad], np.sqrt(ref_d[~pad]))
        assert (d_np[pad] == np.float32(1e5)).all()


def test_exact_wide_k_expands(oracle):
    """k = 96 with cells far too small for it: every query expands."""
    from flashknn_exprepo_amd import synthetic
    rng = np.random.default_rng(18)
    a = synthetic.subsample_to(synthetic.room(seed=48, target=30000), 6000, rng)
    b = synthetic.subsample_to(synthetic.lidar(seed=49), 7000, rng)
    xyz, off = synthetic.pack([a, b])
    ref_i, ref_d = oracle.brute_knn(xyz, off, xyz, off, 96)
    eng = fk.ExactKNN()
    idx, d = eng(_t(xyz), _t(off), 96, cell=torch.tensor([0.02, 0.05], device=_dev()),
                 stats=True)
    _same(idx, d, ref_i, ref_d)
    st = eng.last_stats
    assert st["certified_first_pass"] + st["expanded"] == len(xyz) and st["expanded"] > 0


def test_batched_k_range_errors():
    """Batched k is capped at 64 on the host; a per-cloud k above the
    launch's k is refused on the device (the kernel skips those tiles)."""
    from flashknn_exprepo_amd import synthetic
    from flashknn_exprepo_amd import _lib as F
    clouds, ks = _batch_case(19)
    xyz, off = synthetic.pack(clouds)
    with pytest.raises(ValueError):
        fk.BatchedKNN(_t(xyz), _t(off), [8, 100, 8, 8, 8, 8])
    bk = fk.BatchedKNN(_t(xyz), _t(off), ks, exact=True)
    with pytest.raises(ValueError):
        bk.grid.batched_query(bk.k, bk.out_offsets, bk.out_idx, bk.out_dist, max_k=16)
        bk.grid.status()
    # the error is sticky on the handle until the next build resets it
    with pytest.raises(ValueError):
        bk.grid.status()


@pytest.mark.parametrize("k", [1, 8, 16, 32])
def test_exact_prune_changes_nothing(oracle, k):
    """FK_QUERY_PRUNE (skip neighbour cells beyond the current k-th distance)
    and FK_QUERY_ONE_LANE (64 instead of 32 queries per wave) are speed
    hints: exact results equal brute force with and without them, at cell
    sizes from much smaller to much larger than the k-th distance."""
    from flashknn_exprepo_amd import _lib as F
    from flashknn_exprepo_amd import synthetic
    rng = np.random.default_rng(22)
    a = synthetic.subsample_to(synthetic.room(seed=48, target=30000), 9000, rng)
    b = synthetic.subsample_to(synthetic.lidar(seed=49), 11000, rng)
    dup = np.repeat(a[:40], 3, 0)                       # exact duplicates: ties at d = 0
    xyz, off = synthetic.pack([np.concatenate([a, dup]), b])
    ref_i, ref_d = oracle.brute_knn(xyz, off, xyz, off, k)
    g = fk.Grid()
    for cell in (0.04, 0.2, 0.7):
        inv = torch.tensor([1 / cell, 1 / (2 * cell)], dtype=torch.float32, device=_dev())
        g.build_xyz(_t(xyz), _t(off), inv)
        for prune in (False, True):
            for one in (False, True):
                idx, d = g.query(k, mode=F.FK_MODE_EXACT, pad_mode=F.FK_PAD_MINUS1,
                                 prune=prune, one_lane=one)
                _same(idx, d, ref_i, ref_d)


@pytest.mark.parametrize("n", [6000, 20000, 40000, 100000])
def test_group_sizes_agree(oracle, n):
    """The first pass widens the lane group of a query (2, 4, 8, 16 lanes)
    as the launch gets smaller: at each size the exact rows equal the
    one-lane launch's on every row and brute force on a sample, for k in
    the three K classes that group, with and without pruning."""
    from flashknn_exprepo_amd import _lib as F
    from flashknn_exprepo_amd import synthetic
    rng = np.random.default_rng(n)
    xyz = synthetic.subsample_to(synthetic.room(seed=48), n, rng)
    n = len(xyz)  # the room holds 98k points
    off = np.array([n], np.int32)
    sample = rng.choice(n, 500, replace=False)
    g = fk.Grid()
    for k in (8, 16, 32):
        side, inv = g.auto_cell(_t(xyz), _t(off), k, 1.6)
        g.build_xyz(_t(xyz), _t(off), inv)
        ref_i, ref_d = oracle.brute_knn(xyz, [n], xyz[sample], [500], k)
        for prune in ((False, True) if k <= 16 else (False,)):
            i1, d1 = g.query(k, mode=F.FK_MODE_EXACT, prune=prune, one_lane=True)
            ig, dg = g.query(k, mode=F.FK_MODE_EXACT, prune=prune)
            assert torch.equal(i1, ig) and torch.equal(d1, dg)
            _same(ig[torch.from_numpy(sample).to(ig.device)],
                  dg[torch.from_numpy(sample).to(dg.device)], ref_i, ref_d)


@pytest.mark.parametrize("cloud,k", [("room", 8), ("room", 16), ("room", 32),
                                     ("lidar", 16), ("room250k", 16)])
def test_planned_first_pass_matches_uniform(oracle, cloud, k):
    """Large self-query launches run the balanced first pass (super-tiles of
    64 queries on 1, 2, 4 or 8 lanes per query by their candidate counts):
    every row equals the uniform launch's (FK_QUERY_UNIFORM), in exact and
    27-cell mode, and brute force on a sample."""
    from flashknn_exprepo_amd import _lib as F
    from flashknn_exprepo_amd import synthetic
    xyz = {"room": lambda: synthetic.room(seed=48),
           "lidar": lambda: synthetic.lidar(seed=49),
           "room250k": lambda: synthetic.room(seed=48, target=250000)}[cloud]()
    n = len(xyz)
    e = fk.ExactKNN()
    e.build(_t(xyz), _t(np.array([n], np.int32)), k)
    g = e.grid
    for mode in (F.FK_MODE_EXACT, F.FK_MODE_CELLS):
        kw = dict(mode=mode, prune=g.prune_hint, one_lane=False)
        iu, du = g.query(k, uniform=True, **kw)
        ip, dp = g.query(k, **kw)
        ip2, dp2 = g.query(k, **kw)  # the plan is kept between calls
        g.status()
        assert torch.equal(iu, ip) and torch.equal(du, dp), mode
        assert torch.equal(ip, ip2) and torch.equal(dp, dp2), mode
    sample = np.random.default_rng(k).choice(n, 1000, replace=False)
    ri, rd = oracle.brute_knn(xyz, [n], xyz[sample], [1000], k)
    ie, de = g.query(k, mode=F.FK_MODE_EXACT, prune=g.prune_hint)
    s_t = torch.from_numpy(sample).to(ie.device)
    _same(ie[s_t], de[s_t], ri, rd)


def test_cells_prune_matches_oracle(oracle):
    """In the 27-cell mode on an xyz-built grid, pruning returns the same
    rows as the unpruned scan and the oracle's candidate-set kNN."""
    from flashknn_exprepo_amd import _lib as F
    from flashknn_exprepo_amd import synthetic
    xyz = synthetic.room(seed=48, target=20000)
    g = fk.Grid()
    side = np.float32(0.15)
    inv = torch.tensor([1 / side], dtype=torch.float32, device=_dev())
    g.build_xyz(_t(xyz), _t(np.array([len(xyz)], np.int32)), inv)
    i0, d0 = g.query(16, mode=F.FK_MODE_CELLS, prune=False)
    i1, d1 = g.query(16, mode=F.FK_MODE_CELLS, prune=True)
    assert torch.equal(i0, i1) and torch.equal(d0, d1)
    invf = np.float32(inv.item())
    cells = np.floor((xyz * invf).astype(np.float32)).astype(np.int64)
    ref_i, ref_d = oracle.grid_knn(xyz, cells, None, xyz, cells, None, np.arange(len(xyz)), 16)
    _same(i1, d1, ref_i, ref_d)


def test_exact_cut_radius(oracle):
    rng = np.random.default_rng(12)
    s = rng.random((5000, 3), dtype=np.float32)
    ref_i, ref_d = oracle.brute_knn(s, [5000], s, [5000], 16, cut_r2=0.003)
    idx, d = fk.knn(_t(s), _t(np.array([5000], np.int32)), 16, cut_r2=0.003)
    _same(idx, d, ref_i, ref_d)


def test_exact_full_size_room_properties(oracle):
    """100k room, K=16: exact vs brute force on a 3000-query sample and
    size-independent properties on every row."""
    from flashknn_exprepo_amd import synthetic
    xyz = synthetic.room(seed=48)
    n = len(xyz)
    idx, d = fk.knn(_t(xyz), _t(np.array([n], np.int32)), 16)
    i_np, d_np = idx.cpu().numpy(), d.cpu().numpy()
    assert (i_np[:, 0] == np.arange(n)).all() and (d_np[:, 0] == 0).all()
    assert (np.diff(d_np, axis=1) >= 0).all()
    sample = np.random.default_rng(13).choice(n, 3000, replace=False)
    ref_i, ref_d = oracle.brute_knn(xyz, [n], xyz[sample], [3000], 16)
    assert np.array_equal(i_np[sample], ref_i)
    assert np.array_equal(d_np[sample].view(np.uint32), ref_d.view(np.uint32))


# ------------------------------------------------------------- batched

def _batch_case(seed=14):
    from flashknn_exprepo_amd import synthetic
    rng = np.random.default_rng(seed)
    room = synthetic.room(seed=48, target=30000)
    lid = synthetic.lidar(seed=49)
    sizes = [2000, 9000, 500, 30000, 4000, 17000]
    clouds = [synthetic.subsample_to(room if i % 2 == 0 else lid, n, rng)
              for i, n in enumerate(sizes)]
    ks = [8, 16, 32, 16, 8, 32]
    return clouds, ks


@pytest.mark.parametrize("exact", [True, False])
def test_batched_equals_per_cloud(oracle, exact):
    from flashknn_exprepo_amd import synthetic
    clouds, ks = _batch_case()
    xyz, off = synthetic.pack(clouds)
    bk = fk.BatchedKNN(_t(xyz), _t(off), ks, exact=exact)
    bk.run()
    start = 0
    for c, (cl, k) in enumerate(zip(clouds, ks)):
        bi, bd = bk.per_cloud(c)
        if exact:
            ref_i, ref_d = oracle.brute_knn(cl, [len(cl)], cl, [len(cl)], k)
        else:
            # cells are floor(x * inv) with the build's fp32 product
            inv = np.float32(1.0 / np.float64(np.float32(bk.cell[c].item())))
            g = np.floor((cl * inv).astype(np.float32)).astype(np.int64)
            ref_i, ref_d = oracle.grid_knn(cl, g, None, cl, g, None, np.arange(len(cl)), k)
        _same(bi - start, bd, ref_i, ref_d)
        start += len(cl)


def test_batched_graph_replay_matches_eager():
    from flashknn_exprepo_amd import synthetic
    clouds, ks = _batch_case(15)
    xyz, off = synthetic.pack(clouds)
    bk = fk.BatchedKNN(_t(xyz), _t(off), ks, exact=True)
    ei, ed = [t.clone() for t in bk.run()]
    bk.capture()
    for rep in range(4):
        bk.out_idx.fill_(-7)
        bk.out_dist.fill_(-7.0)
        torch.cuda.synchronize()
        ri, rd = bk.replay()
        torch.cuda.synchronize()
        assert torch.equal(ri, ei) and torch.equal(rd, ed), f"replay {rep}"
        _ = (torch.randn(1000, device=_dev()) * 2).sum().item()  # eager work between


def test_graph_replay_refused_after_rebuild():
    """A captured step holds the build's pointers and sizes: after a rebuild
    (or a reallocation) the replay is refused on the host and launches
    nothing; a new capture serves again."""
    from flashknn_exprepo_amd import _lib as F
    from flashknn_exprepo_amd import synthetic
    clouds, ks = _batch_case(16)
    xyz, off = synthetic.pack(clouds)
    bk = fk.BatchedKNN(_t(xyz), _t(off), ks, exact=True)
    ei, ed = [t.clone() for t in bk.run()]
    bk.capture()
    bk.replay()
    torch.cuda.synchronize()
    bk.grid.build_xyz(bk.xyz, bk.offsets, (1.0 / bk.cell.double()).float())
    bk.out_idx.fill_(-7)
    with pytest.raises(fk.FlashKNNError) as err:
        bk.grid.replay()
    assert err.value.status == F.FK_ERR_STATE
    torch.cuda.synchronize()
    assert (bk.out_idx == -7).all()
    bk.capture()
    ri, rd = bk.replay()
    torch.cuda.synchronize()
    assert torch.equal(ri, ei) and torch.equal(rd, ed)


def test_failed_build_leaves_the_handle_unbuilt(oracle):
    """A build that returns an error leaves no usable build behind (it may
    already have reallocated tables that the earlier build's sizes
    describe): queries and attaches are refused with FK_ERR_STATE until a
    build succeeds, and that build queries exactly."""
    from flashknn_exprepo_amd import _lib as F
    from flashknn_exprepo_amd.grid import Grid
    rng = np.random.default_rng(73)
    n = 3000
    grid = rng.integers(0, 6, (n, 3))
    g = Grid()
    g.reserve(n, 1 << 16, 4, n)
    g.build_cells(_t(grid), None, None, num_clouds=1, no_sync=True)
    g.status()
    g.query(8)
    with pytest.raises(fk.FlashKNNError) as err:  # more points than reserved
        g.build_cells(_t(rng.integers(0, 6, (2 * n, 3))), None, None, num_clouds=1,
                      no_sync=True)
    assert err.value.status == F.FK_ERR_CAPACITY
    with pytest.raises(fk.FlashKNNError) as err:
        g.query(8)
    assert err.value.status == F.FK_ERR_STATE
    with pytest.raises(fk.FlashKNNError) as err:
        g.attach_queries(_t(grid.astype(np.float32)), qcell=_t(grid))
    assert err.value.status == F.FK_ERR_STATE
    g.build_cells(_t(grid), None, None, num_clouds=1)
    idx, d = g.query(8)
    ref_i, ref_d = oracle.flashknn_query(grid, None, None, 8, 0)
    _same(idx, d, ref_i, ref_d)
    # a reserve that reallocates the build's tables also leaves it unbuilt
    g.reserve(4 * n, 1 << 18, 4, 4 * n)
    with pytest.raises(fk.FlashKNNError) as err:
        g.query(8)
    assert err.value.status == F.FK_ERR_STATE
    g.build_cells(_t(grid), None, None, num_clouds=1, no_sync=True)
    g.status()
    idx, d = g.query(8)
    _same(idx, d, ref_i, ref_d)


def test_batched_query_after_overflowing_no_sync_build():
    """A no-sync build whose boxes outgrow the reserve still records every
    cloud's base, past the cell table: a batched query on it reads no cell
    table and writes nothing, the next status read reports the overflow, and a
    synchronising rebuild serves the same handle again."""
    from flashknn_exprepo_amd import _lib as F
    from flashknn_exprepo_amd import synthetic
    clouds, ks = _batch_case(17)
    xyz, off = synthetic.pack(clouds)
    bk = fk.BatchedKNN(_t(xyz), _t(off), ks, exact=True)
    ei, ed = [t.clone() for t in bk.run()]
    g = bk.grid
    g.reserve(len(xyz), 1 << 16, len(ks), len(xyz))
    fine = torch.full((len(ks),), 1000.0, dtype=torch.float32, device=_dev())  # 1 mm cells
    g.build_xyz(bk.xyz, bk.offsets, fine, no_sync=True)
    bk.out_idx.fill_(-7)
    bk.out_dist.fill_(-7.0)
    bk.run()
    with pytest.raises(fk.FlashKNNError) as err:
        g.status()
    assert err.value.status == F.FK_ERR_GRID_TOO_LARGE
    assert (bk.out_idx == -7).all() and (bk.out_dist == -7.0).all()
    g.build_xyz(bk.xyz, bk.offsets, (1.0 / bk.cell.double()).float())
    ri, rd = bk.run()
    g.status()
    assert torch.equal(ri, ei) and torch.equal(rd, ed)


# ------------------------------------------------------------ backward

def test_dist_backward_vs_oracle(oracle):
    rng = np.random.default_rng(16)
    p = rng.random((2000, 3), dtype=np.float32)
    q = rng.random((700, 3), dtype=np.float32)
    idx, _ = oracle.brute_knn(p, [2000], q, [700], 16)
    idx[3, 5] = -1
    g = rng.standard_normal((700, 16)).astype(np.float32)
    qt = _t(q).requires_grad_(True)
    pt = _t(p).requires_grad_(True)
    d = fk.gathered_sqdist(qt, pt, _t(idx))
    d = torch.where(_t(idx) >= 0, d, torch.zeros_like(d))
    (d * _t(g)).sum().backward()
    gq, gp = oracle.dist_backward(q, p, idx, g)
    for got, ref in ((qt.grad, gq), (pt.grad, gp)):
        got = got.double().cpu().numpy()
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert rel <= 1e-5, rel


def test_backward_deterministic():
    rng = np.random.default_rng(17)
    p = _t(rng.random((3000, 3), dtype=np.float32))
    q = _t(rng.random((3000, 3), dtype=np.float32))
    idx, _ = fk.knn(p, _t(np.array([3000], np.int32)), 16, q,
                    _t(np.array([3000], np.int32)))
    feat = _t(rng.standard_normal((3000, 64)).astype(np.float32))
    outs = []
    for _ in range(3):
        pp, qq, ff = (t.clone().requires_grad_(True) for t in (p, q, feat))
        out = fk.idw_interpolate(ff, qq, pp, idx)
        out.square().sum().backward()
        outs.append([t.grad.clone() for t in (pp, qq, ff)])
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert torch.equal(a, b)


def test_idw_interp_padding_and_channels(oracle):
    """Fused IDW forward/backward with padded slots (idx = -1), a fully
    padded row and 70 channels (two 64-lane chunks) against an FP64 torch
    reference on the CPU."""
    rng = np.random.default_rng(23)
    p = rng.random((500, 3), dtype=np.float32)
    q = rng.random((300, 3), dtype=np.float32)
    idx, _ = oracle.brute_knn(p, [500], q, [300], 8)
    idx[5, 3:] = -1
    idx[7, :] = -1
    feat = rng.standard_normal((500, 70)).astype(np.float32)
    wout = rng.standard_normal((300, 70)).astype(np.float32)
    eps = 1e-6
    qt, pt, ft = (_t(a).requires_grad_(True) for a in (q, p, feat))
    out = fk.idw_interpolate(ft, qt, pt, _t(idx), eps=eps)
    (out * _t(wout)).sum().backward()
    it = torch.from_numpy(idx)
    q_, p_, f_ = (torch.from_numpy(a.astype(np.float64)).requires_grad_(True)
                  for a in (q, p, feat))
    ok = it >= 0
    safe = it.clamp_min(0)
    d_ = ((q_[:, None, :] - p_[safe]) ** 2).sum(-1)
    w_ = torch.where(ok, 1.0 / (d_ + eps), torch.zeros_like(d_))
    S = w_.sum(1, keepdim=True)
    wn_ = torch.where(S > 0, w_ / S.clamp_min(1e-300), torch.zeros_like(w_))
    out_ = (wn_[..., None] * f_[safe]).sum(1)
    (out_ * torch.from_numpy(wout.astype(np.float64))).sum().backward()
    got = out.detach().double().cpu()
    assert torch.all(got[7] == 0)
    assert (torch.linalg.norm(got - out_.detach()) / torch.linalg.norm(out_.detach())) <= 1e-5
    for g, r in ((qt.grad, q_.grad), (pt.grad, p_.grad), (ft.grad, f_.grad)):
        g = g.double().cpu()
        assert (torch.linalg.norm(g - r) / torch.linalg.norm(r)) <= 1e-5
    assert torch.all(qt.grad[7] == 0)


def test_idw_fp64_finite_differences(oracle):
    """Backward check on a 2k-point copy: gradients of the fp32 path vs FP64
    central differences of the same loss, relative error <= 1e-5."""
    from flashknn_exprepo_amd import synthetic
    fine = synthetic.room(seed=51, target=20000)
    fine = synthetic.subsample_to(fine, 2000, np.random.default_rng(0))
    coarse = fine[synthetic.voxel_first(fine, 0.16)]
    idx, _ = oracle.brute_knn(coarse, [len(coarse)], fine, [len(fine)], 16)
    rng = np.random.default_rng(18)
    feat = rng.standard_normal((len(coarse), 8))
    wout = rng.standard_normal((len(fine), 8))
    eps = 1e-4

    def loss64(qx, px, f):
        d = ((qx[:, None, :] - px[idx]) ** 2).sum(-1)
        w = 1.0 / (d + eps)
        wn = w / w.sum(1, keepdims=True)
        return ((wn[..., None] * f[idx]).sum(1) * wout).sum()

    qt = _t(fine).requires_grad_(True)
    pt = _t(coarse).requires_grad_(True)
    ft = _t(feat.astype(np.float32)).requires_grad_(True)
    out = fk.idw_interpolate(ft, qt, pt, _t(idx), eps=eps)
    (out * _t(wout.astype(np.float32))).sum().backward()
    q64, p64 = fine.astype(np.float64), coarse.astype(np.float64)
    # analytic fp64 gradient through torch autograd on the CPU (independent)
    q_ = torch.from_numpy(q64).requires_grad_(True)
    p_ = torch.from_numpy(p64).requires_grad_(True)
    f_ = torch.from_numpy(feat).requires_grad_(True)
    it = torch.from_numpy(idx)
    d_ = ((q_[:, None, :] - p_[it]) ** 2).sum(-1)
    w_ = 1.0 / (d_ + eps)
    wn_ = w_ / w_.sum(1, keepdim=True)
    ((wn_[..., None] * f_[it]).sum(1) * torch.from_numpy(wout)).sum().backward()
    for got, ref in ((qt.grad, q_.grad), (pt.grad, p_.grad), (ft.grad, f_.grad)):
        got = got.double().cpu().numpy()
        ref = ref.numpy()
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert rel <= 1e-5, rel
    # FP64 central differences on a few coordinates pin the fp64 autograd
    h = 1e-6
    for (arr, grad) in ((q64, q_.grad.numpy()), (p64, p_.grad.numpy())):
        for i in range(0, len(arr), max(1, len(arr) // 5)):
            for a in range(3):
                old = arr[i, a]
                arr[i, a] = old + h
                lp = loss64(q64, p64, feat)
                arr[i, a] = old - h
                lm = loss64(q64, p64, feat)
                arr[i, a] = old
                num = (lp - lm) / (2 * h)
                assert abs(num - grad[i, a]) <= 1e-4 * max(1.0, abs(num)), (i, a, num, grad[i, a])


def test_pointops_interpolation2_matches_torch():
    rng = np.random.default_rng(19)
    s = _t(rng.random((2000, 3), dtype=np.float32))
    q = _t(rng.random((3000, 3), dtype=np.float32))
    so, qo = _t(np.array([2000], np.int32)), _t(np.array([3000], np.int32))
    feat = _t(rng.standard_normal((2000, 32)).astype(np.float32)).requires_grad_(True)
    out = fk.pointops.interpolation2(s, q, feat, so, qo, 3)
    ref = fk.pointops.interpolation(s, q, feat.detach(), so, qo, 3)
    torch.testing.assert_close(out, ref, rtol=1e-6, atol=1e-6)
    g = torch.randn_like(out)
    out.backward(g)
    idx, dist = fk.pointops.knn_query(3, s, so, q, qo)
    w = 1.0 / (dist.double() + 1e-8)
    w = w / w.sum(1, keepdim=True)
    ref_g = torch.zeros(2000, 32, dtype=torch.float64, device=_dev())
    ref_g.index_add_(0, idx.long().reshape(-1),
                     (w[..., None] * g.double()[:, None, :]).reshape(-1, 32))
    torch.testing.assert_close(feat.grad.double(), ref_g, rtol=1e-5, atol=1e-6)


# --------------------------------------------------------------- misc

def test_morton_kernel_matches_oracle(oracle):
    rng = np.random.default_rng(20)
    g = rng.integers(-5, 1 << 16, (5000, 3))
    b = rng.integers(0, 100, 5000)
    k = fk.xyz2key(_t(g[:, 0]), _t(g[:, 1]), _t(g[:, 2]), _t(b))
    assert np.array_equal(k.cpu().numpy(), oracle.morton_encode(g, b))
    gd = rng.integers(0, 1000, (300, 3))
    nb = fk.functions.neighbor_keys(_t(gd), _t(b[:300]))
    assert np.array_equal(nb.cpu().numpy(), oracle.morton_neighbor_keys(gd, b[:300]))


def test_grid_tables_partition(oracle):
    xyz = np.random.default_rng(21).random((7000, 3), dtype=np.float32)
    g = fk.Grid()
    g.build_xyz(_t(xyz), _t(np.array([7000], np.int32)), _t(np.array([10.0], np.float32)))
    rec, cs = g.tables()
    rec, cs = rec.cpu().numpy(), cs.cpu().numpy()
    ids = rec[:, 3].view(np.int32)
    assert np.array_equal(np.sort(ids), np.arange(7000))
    assert (np.diff(cs) >= 0).all() and cs[-1] == 7000
    cells = np.floor(xyz * np.float32(10.0)).astype(np.int64)
    lin = (cells[:, 0] * 10 + cells[:, 1]) * 10 + cells[:, 2]
    counts = np.bincount(lin, minlength=1000)
    assert np.array_equal(np.diff(cs), counts)
    assert np.array_equal(np.sort(lin[ids]), lin[ids])  # records are cell-sorted


def test_auto_cell_matches_host_formula():
    from flashknn_exprepo_amd.knn import _cell_sizes_torch, cell_sizes
    rng = np.random.default_rng(22)
    a = rng.random((3000, 3), dtype=np.float32)
    b = (rng.random((5000, 3), dtype=np.float32) * [10, 10, 0.001]).astype(np.float32)
    xyz, off = np.concatenate([a, b]), np.array([3000, 8000], np.int32)
    gpu = cell_sizes(_t(xyz), _t(off), 16, scale=1.0).cpu()
    cpu = _cell_sizes_torch(torch.from_numpy(xyz), torch.from_numpy(off), 16, 1.0)
    torch.testing.assert_close(gpu, cpu, rtol=1e-5, atol=0)


def test_exact_far_queries_expand_to_cover(oracle):
    rng = np.random.default_rng(23)
    s = rng.random((2000, 3), dtype=np.float32)
    q = np.array([[50, 50, 50], [-30, 0.5, 0.5], [0.5, 0.5, 0.5], [1.2, -0.3, 0.7]],
                 np.float32)
    ref_i, ref_d = oracle.brute_knn(s, [2000], q, [4], 16)
    eng = fk.ExactKNN()
    idx, d = eng(_t(s), _t(np.array([2000], np.int32)), 16, _t(q),
                 _t(np.array([4], np.int32)), stats=True)
    _same(idx, d, ref_i, ref_d)
    assert eng.last_stats["expanded"] >= 2


def test_exact_sparse_clusters(oracle):
    """Dense blobs far apart: most blocks hold fewer than k points, so the
    expansion has to double its radius before it can certify."""
    rng = np.random.default_rng(24)
    blobs = [rng.normal(c, 0.01, (20, 3)) for c in rng.uniform(-20, 20, (40, 3))]
    s = np.concatenate(blobs).astype(np.float32)
    ref_i, ref_d = oracle.brute_knn(s, [len(s)], s, [len(s)], 32)
    idx, d = fk.knn(_t(s), _t(np.array([len(s)], np.int32)), 32, cell=torch.tensor([0.5]))
    _same(idx, d, ref_i, ref_d)


def test_exact_build_explicit_cell_beyond_reserve(oracle):
    """ExactKNN.build with a caller's cell whose dense box (100^3 cells)
    exceeds the no-sync reserve of 3,000 points: the build sizes its box
    exactly by default, so querying the grid directly is exact."""
    from flashknn_exprepo_amd import _lib as F
    s = (np.random.default_rng(30).random((3000, 3)) * 20).astype(np.float32)
    ref_i, ref_d = oracle.brute_knn(s, [3000], s, [3000], 8)
    eng = fk.ExactKNN()
    eng.build(_t(s), _t(np.array([3000], np.int32)), 8, cell=torch.tensor([0.2]))
    idx, d = eng.grid.query(8, mode=F.FK_MODE_EXACT, pad_mode=F.FK_PAD_MINUS1)
    eng.grid.status()
    _same(idx, d, ref_i, ref_d)


@pytest.mark.parametrize("columns", [False, True])
@pytest.mark.parametrize("k", [4, 16, 32])
def test_exact_isolated_queries_radius_probe(oracle, k, columns):
    """Queries short of k candidates after the first pass take their first
    expansion shell from the full-height block-count probe: isolated points
    2-14 m from a dense floor patch (jumps of up to 31 cells), points in
    the box corners (probe radii clipped to the cover radius), and a tall
    stack of points above an isolated query (the full-height count then
    overstates the block, so the growth goes on after the jump); exact
    against brute force on the dense and the column layout."""
    from flashknn_exprepo_amd import _lib as F
    rng = np.random.default_rng(31)
    floor = np.c_[rng.uniform(0, 4, (3000, 2)), rng.normal(0, 0.01, 3000)]
    far = np.c_[rng.uniform(-14, 18, (60, 2)), rng.uniform(-0.5, 3, 60)]
    corners = np.array([[-14, -14, -0.5], [18, 18, 3], [18, -14, 3], [-14, 18, -0.5]])
    stack = np.c_[np.full((40, 2), [9.0, -6.0]) + rng.normal(0, 0.02, (40, 2)),
                  np.linspace(0.5, 3.0, 40)]
    s = np.concatenate([floor, far, corners, stack, [[9.2, -6.1, -0.4]]]).astype(np.float32)
    n = len(s)
    ref_i, ref_d = oracle.brute_knn(s, [n], s, [n], k)
    from flashknn_exprepo_amd.grid import Grid
    g = Grid().build_xyz(_t(s), _t(np.array([n], np.int32)),
                         torch.tensor([1 / 0.1], dtype=torch.float32, device=_dev()),
                         columns=columns)
    idx, d = g.query(k, mode=F.FK_MODE_EXACT, pad_mode=F.FK_PAD_MINUS1)
    g.status()
    _same(idx, d, ref_i, ref_d)
    assert g.query_stats()["expanded"] >= 60


@pytest.mark.parametrize("k", [8, 16, 100])
def test_exact_culled_shells_far_offset(oracle, k):
    """Expansion shells are culled against the current k-th distance: dense
    blobs plus sparse points far from them, the cloud shifted thousands of
    metres from the origin (large scaled coordinates), fine cells so that the
    sparse queries grow shells of many cells; exact against brute force."""
    rng = np.random.default_rng(29)
    blobs = [rng.normal(c, 0.05, (400, 3)) for c in rng.uniform(-3, 3, (6, 3))]
    sparse = rng.uniform(-12, 12, (300, 3))
    s = (np.concatenate(blobs + [sparse]) + np.array([4096.0, -2500.0, 731.0])).astype(np.float32)
    ref_i, ref_d = oracle.brute_knn(s, [len(s)], s, [len(s)], k)
    for cell in (0.05, 0.2):
        idx, d = fk.knn(_t(s), _t(np.array([len(s)], np.int32)), k, cell=torch.tensor([cell]))
        _same(idx, d, ref_i, ref_d)


@pytest.mark.parametrize("k", [8, 16])
def test_adaptive_neighborhood_lidar(oracle, k):
    from flashknn_exprepo_amd import synthetic
    rng = np.random.default_rng(25)
    pts = synthetic.subsample_to(synthetic.lidar(seed=49), 20000, rng)
    grid = np.floor(pts / np.float32(0.05)).astype(np.int64)
    batch = np.zeros(len(pts), np.int64)
    ref_i, ref_d, ref_sel = oracle.adaptive_query(grid, batch, pts, k)
    a = fk.AdaptiveNeighborhoodFlashKNN(num_nbr=k)
    idx, d = a.query(_t(grid), _t(batch), _t(pts), return_distances=True)
    _same(idx, d, ref_i, ref_d)
    assert sum(s["query_points"] for s in a.last_stats["selected_levels"]) == len(pts)


def test_adaptive_neighborhood_batches_no_xyz(oracle):
    rng = np.random.default_rng(26)
    g1 = rng.integers(0, 30, (3000, 3))
    g2 = rng.integers(100, 400, (2000, 3))
    grid = np.concatenate([g1, g2]).astype(np.int64)
    batch = np.concatenate([np.zeros(3000), np.ones(2000)]).astype(np.int64)
    ref_i, ref_d, _ = oracle.adaptive_query(grid, batch, None, 12)
    a = fk.AdaptiveNeighborhoodFlashKNN(num_nbr=12)
    idx, d = a.query(_t(grid), _t(batch), None, return_distances=True)
    _same(idx, d, ref_i, ref_d)
    sel = fk.select_adaptive_levels(fk.build_adaptive_octree(_t(grid), _t(batch)), 12)
    assert sel.dtype == torch.int16


def test_adaptive_neighborhood_interleaved_batches(oracle):
    """Three batches stored interleaved (every chunk of the per-batch minimum
    mixes batches) with offsets far from zero."""
    rng = np.random.default_rng(27)
    n = 4500
    batch = rng.integers(0, 3, n).astype(np.int64)
    grid = (rng.integers(0, 200, (n, 3)) + np.array([[-10**9, 5 * 10**8, 7]])
            + 1000 * batch[:, None]).astype(np.int64)
    ref_i, ref_d, _ = oracle.adaptive_query(grid, batch, None, 8)
    a = fk.AdaptiveNeighborhoodFlashKNN(num_nbr=8)
    idx, d = a.query(_t(grid), _t(batch), None, return_distances=True)
    _same(idx, d, ref_i, ref_d)


@pytest.mark.parametrize("n,B,contiguous", [(1, 1, True), (3000, 1, True), (5000, 4, True),
                                            (5000, 4, False), (70000, 9, True)])
def test_batch_min(n, B, contiguous):
    from flashknn_exprepo_amd.functions.adaptive_octree import batch_min
    rng = np.random.default_rng(n + B)
    b = rng.integers(0, B, n)
    if contiguous:
        b.sort()
    v = rng.integers(-10**12, 10**12, (n, 3))
    ref = np.full((B, 3), np.iinfo(np.int64).max, np.int64)
    np.minimum.at(ref, b, v)
    out = batch_min(_t(v.astype(np.int64)), _t(b.astype(np.int64)), B)
    assert np.array_equal(out.cpu().numpy(), ref)
