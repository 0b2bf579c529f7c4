"""Voxel-first subsampling and the level hierarchy on the device against the
oracle (needs an MI355X).  Indices are compared exactly."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

fk = pytest.importorskip("flashknn_exprepo_amd")


def _dev():
    return torch.device("cuda", 0)


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(_dev())


@pytest.mark.parametrize("size", [0.06, 0.04, 0.1, 0.12 * 2])
def test_cell_quotient_is_correctly_rounded(size):
    """Cells of the back query and the hierarchy levels come from the
    correctly rounded fp32 quotient x / size (what the oracle restates), on
    values where the product with fl32(1/size) floors differently."""
    from flashknn_exprepo_amd.functions.FlashKnnWrapper import quotient32
    rng = np.random.default_rng(20)
    x = np.concatenate([np.array([5.0399995, 7.5599995, 3.3599997, 1.5999999, 2.1,
                                  -5.0399995, 0.0, -0.0], np.float32),
                        (rng.random(1 << 20, dtype=np.float32) * 16 - 8).astype(np.float32)])
    got = quotient32(_t(x), size).cpu().numpy()
    assert np.array_equal(got.view(np.uint32), (x / np.float32(size)).view(np.uint32))


@pytest.mark.parametrize("size", [0.04, 0.06, 0.12, 0.24, 0.48])
def test_voxel_first_room(oracle, size):
    from flashknn_exprepo_amd import synthetic
    xyz = synthetic.room(seed=48)
    got = fk.voxel_first(_t(xyz), size).cpu().numpy()
    assert np.array_equal(got, oracle.voxel_first(xyz, size))


def test_voxel_first_batched_lidar_negative(oracle):
    from flashknn_exprepo_amd import synthetic
    rng = np.random.default_rng(21)
    a = synthetic.lidar(seed=49)                       # negative coordinates
    b = synthetic.subsample_to(synthetic.room(seed=48, target=30000), 20000, rng)
    xyz, off = synthetic.pack([a, b, a[:500]])
    got = fk.voxel_first(_t(xyz), 0.2, _t(off)).cpu().numpy()
    assert np.array_equal(got, oracle.voxel_first(xyz, 0.2, off))


def test_voxel_first_aliasing_and_edges(oracle):
    a = np.array([[0.5, 0, 0], [65536.5, 0, 0], [-0.5, 0, 0], [65535.5, 0, 0],
                  [0.5, 1.5, 0]], np.float32)
    assert fk.voxel_first(_t(a), 1.0).tolist() == [0, 2, 4]
    assert fk.voxel_first(_t(a[:1]), 1.0).tolist() == [0]
    assert fk.voxel_first(_t(np.zeros((0, 3), np.float32)), 1.0).numel() == 0
    dup = np.repeat(a[:1], 5000, 0)
    assert fk.voxel_first(_t(dup), 0.01).tolist() == [0]
    bad = a.copy()
    bad[3, 1] = np.inf
    with pytest.raises(ValueError):
        fk.voxel_first(_t(bad), 1.0)
    with pytest.raises(ValueError):
        fk.voxel_first(_t(a), 0.0)
    # a call after the device-side error starts clean
    assert fk.voxel_first(_t(a), 1.0).tolist() == [0, 2, 4]


def _oracle_hierarchy(oracle, xyz, sizes, ks, alpha):
    e = int(round(np.log2(alpha)))
    cur, cum = xyz, np.arange(len(xyz))
    fwd, levels = [], []
    for level, (size, k) in enumerate(zip(sizes, ks)):
        if level:
            local = oracle.voxel_first(cur, size)
            cur, cum = cur[local], cum[local]
            fwd.append(local)
            levels.append(cum)
        grid = np.floor(cur / np.float32(size)).astype(np.int64)
        fwd.append(oracle.flashknn_query(grid, None, cur, k, e)[0])
    back = []
    for level in range(len(levels) - 1, -1, -1):
        s = sizes[level + 1]
        back.append(oracle.back_query(xyz, levels[level], s * 2, None, 1)[0][:, 0])
    return fwd + back


@pytest.mark.parametrize("alpha,k", [(2, 16), (4, 16), (16, 24)])
def test_build_flash_hierarchy_vs_oracle(oracle, alpha, k):
    """Networks/hierarchy.py:39-97 end to end on the device: voxel-first
    levels, FlashKNN.query per level, back queries to level 0."""
    from flashknn_exprepo_amd import synthetic
    xyz = synthetic.room(seed=48, target=20000)
    sizes, ks = (0.06, 0.12, 0.24, 0.48), (k,) * 4
    got = fk.build_flash_hierarchy(_t(xyz), sizes, ks, alpha)
    ref = _oracle_hierarchy(oracle, xyz, sizes, ks, alpha)
    assert len(got) == len(ref) == 10
    for j, (g, r) in enumerate(zip(got, ref)):
        assert np.array_equal(g.cpu().numpy(), r), f"output {j}"
