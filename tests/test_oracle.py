"""Pins the CPU oracle (CPU only).

The reference ships no per-query golden vectors (its sample room is not in
the repository), so the oracle is pinned by closed forms, hand vectors,
independent restatements (bit-loop Morton, pure-Python candidate
enumeration, numpy fp32 brute force, scipy's KD-tree in float64) and
invariants chosen so that a dropped term, a wrong tie-break, a non-strict
cut or a wrong padding value fails at least one test.
"""
import numpy as np
import pytest


# ------------------------------------------------------------------ Morton

@pytest.mark.parametrize("xyz,key", [
    ((1, 0, 0), 4), ((0, 1, 0), 2), ((0, 0, 1), 1), ((1, 1, 1), 7),
    ((2, 0, 0), 32), ((0, 2, 0), 16), ((0, 0, 2), 8), ((3, 0, 0), 36),
    ((0xFFFF, 0, 0), int("100" * 16, 2)),
])
def test_morton_hand_vectors(oracle, xyz, key):
    assert oracle.morton_encode(np.array([xyz]))[0] == key
    assert oracle.morton_key_bitloop(*xyz) == key


def test_morton_batch_bits(oracle):
    g = np.array([[0, 0, 0], [1, 1, 1], [5, 9, 3]])
    b = np.array([1, 7, 32767])
    k = oracle.morton_encode(g, b)
    assert k[0] == 1 << 48
    assert k[1] == (7 << 48) | 7
    assert k[2] >> 48 == 32767


def test_morton_random_vs_bitloop(oracle):
    rng = np.random.default_rng(0)
    g = rng.integers(0, 1 << 16, size=(500, 3))
    b = rng.integers(0, 1 << 15, size=500)
    k = oracle.morton_encode(g, b)
    for i in range(500):
        assert k[i] == oracle.morton_key_bitloop(*map(int, g[i]), int(b[i]))


def test_morton_masks_16_bits(oracle):
    # construction.cu:15-17 keeps the low 16 bits of each coordinate.
    a = oracle.morton_encode(np.array([[1 << 16 | 5, 3, 2]]))
    b = oracle.morton_encode(np.array([[5, 3, 2]]))
    assert a[0] == b[0]


def test_neighbor_keys_order_centre_first(oracle):
    g = np.array([[5, 6, 7]])
    k = oracle.morton_neighbor_keys(g, np.array([2]))
    offs = [0, -1, 1]
    exp = [oracle.morton_key_bitloop(5 + offs[n // 9], 6 + offs[(n // 3) % 3],
                                     7 + offs[n % 3], 2) for n in range(27)]
    assert list(k) == exp
    assert k[0] == oracle.morton_encode(g, np.array([2]))[0]


# -------------------------------------------------------- brute-force kNN

def _cloud(rng, n, scale=1.0):
    return (rng.random((n, 3), dtype=np.float32) * scale).astype(np.float32)


def test_brute_c_vs_numpy(oracle):
    rng = np.random.default_rng(1)
    s = _cloud(rng, 700)
    q = _cloud(rng, 120)
    i_c, d_c = oracle.brute_knn(s, [700], q, [120], 16)
    i_n, d_n = oracle.brute_knn_numpy(s, q, 16)
    assert np.array_equal(i_c, i_n)
    assert np.array_equal(d_c.view(np.uint32), d_n.view(np.uint32))


def test_brute_vs_scipy_kdtree(oracle):
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(2)
    s = _cloud(rng, 3000)
    i_c, d_c = oracle.brute_knn(s, [3000], s, [3000], 8, f64=True)
    dd, ii = cKDTree(s.astype(np.float64)).query(s.astype(np.float64), 8)
    np.testing.assert_allclose(np.sqrt(d_c), dd, rtol=1e-12, atol=1e-12)
    # indices agree wherever there is no exact fp64 tie
    gap = np.diff(d_c, axis=1).min(1) > 0
    assert np.array_equal(i_c[gap], ii[gap])


def test_brute_fp32_vs_fp64(oracle):
    rng = np.random.default_rng(3)
    s = _cloud(rng, 2000)
    i32, d32 = oracle.brute_knn(s, [2000], s, [2000], 16)
    i64, d64 = oracle.brute_knn(s, [2000], s, [2000], 16, f64=True)
    np.testing.assert_allclose(d32, d64, rtol=2e-6, atol=1e-12)
    # sets agree except where fp32 rounding reorders a near tie at the K-th
    same = (np.sort(i32, 1) == np.sort(i64, 1)).all(1)
    assert same.mean() > 0.99


def test_brute_segments_and_padding(oracle):
    rng = np.random.default_rng(4)
    s = _cloud(rng, 30)
    q = _cloud(rng, 9)
    soff, qoff = [5, 25, 30], [3, 6, 9]
    idx, d = oracle.brute_knn(s, soff, q, qoff, 8)
    # segment 0 holds 5 support points: slots 5..7 are padding (-1, inf)
    assert (idx[:3, 5:] == -1).all() and np.isinf(d[:3, 5:]).all()
    assert ((idx[:3, :5] >= 0) & (idx[:3, :5] < 5)).all()
    assert ((idx[3:6] >= 5) & (idx[3:6] < 25)).all()
    assert ((idx[6:, :5] >= 25) & (idx[6:, :5] < 30)).all()
    ref_i, ref_d = oracle.brute_knn_numpy(s[5:25], q[3:6], 8)
    assert np.array_equal(idx[3:6], ref_i + 5)


def test_brute_ties_smaller_index_first(oracle):
    # duplicated points: equal distances must come out in index order
    s = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 0], [0, 1, 0], [1, 0, 0]], np.float32)
    q = np.array([[0, 0, 0]], np.float32)
    idx, d = oracle.brute_knn(s, [5], q, [1], 5)
    assert list(idx[0]) == [0, 1, 2, 3, 4]
    s2 = s[[4, 3, 2, 1, 0]]
    idx2, _ = oracle.brute_knn(s2, [5], q, [1], 5)
    assert list(idx2[0]) == [4, 0, 1, 2, 3]


def test_brute_cut_is_strict(oracle):
    s = np.array([[0, 0, 0], [0.5, 0, 0], [1, 0, 0], [2, 0, 0]], np.float32)
    q = np.array([[0, 0, 0]], np.float32)
    idx, d = oracle.brute_knn(s, [4], q, [1], 4, cut_r2=1.0)
    assert list(idx[0]) == [0, 1, -1, -1]


def test_distance_expression_order(oracle):
    # ((xx+yy)+zz) differs from (xx+(yy+zz)) for these values: pins the order.
    f = np.float32
    a = f(2.45e-4)
    p = np.array([[f(1.0), a, a]], np.float32)
    q = np.zeros((1, 3), np.float32)
    _, d = oracle.brute_knn(p, [1], q, [1], 1)
    yy = f(a * a)
    left = f(f(f(1.0) + yy) + yy)
    right = f(f(1.0) + f(yy + yy))
    assert left != right
    assert d[0, 0] == left


# -------------------------------------------------- grid candidate kNN

def _grid_case(rng, n, nb, span):
    xyz = _cloud(rng, n, span)
    cell = np.floor(xyz / np.float32(1.0)).astype(np.int64)
    batch = rng.integers(0, nb, n)
    return xyz, cell, batch


@pytest.mark.parametrize("seed,R,cut", [(0, 1, np.inf), (1, 1, 0.7), (2, 2, np.inf)])
def test_grid_c_vs_python_enumeration(oracle, seed, R, cut):
    rng = np.random.default_rng(seed)
    xyz, cell, batch = _grid_case(rng, 160, 2, 5.0)
    xyz[7] = xyz[3]          # exact duplicate: a tie
    cell[7] = cell[3]
    batch[7] = batch[3]
    q = rng.choice(160, 40, replace=False)
    pad = q.copy()
    a_i, a_d = oracle.grid_knn(xyz, cell, batch, xyz[q], cell[q], batch[q], pad, 12,
                               R, cut * cut)
    b_i, b_d = oracle.grid_knn_python(xyz, cell, batch, xyz[q], cell[q], batch[q],
                                      pad, 12, R, cut * cut)
    assert np.array_equal(a_i, b_i)
    assert np.array_equal(a_d.view(np.uint32), b_d.view(np.uint32))


def test_grid_equals_brute_when_block_covers_cloud(oracle):
    rng = np.random.default_rng(5)
    xyz = _cloud(rng, 900, 3.0)
    cell = np.floor(xyz).astype(np.int64)        # 3x3x3 cells: R=1 from the
    q = np.where((cell == 1).all(1))[0]          # centre cell covers all
    gi, gd = oracle.grid_knn(xyz, cell, None, xyz[q], cell[q], None, q, 16, 1)
    bi, bd = oracle.brute_knn(xyz, [900], xyz[q], [len(q)], 16)
    assert len(q) > 0
    assert np.array_equal(gi, bi) and np.array_equal(gd, bd)


def test_grid_padding_is_self_and_inf(oracle):
    xyz = np.array([[0.1, 0.1, 0.1], [0.2, 0.1, 0.1], [9.5, 9.5, 9.5]], np.float32)
    g = np.floor(xyz / 1.0).astype(np.int64)
    idx, d = oracle.flashknn_query(g, None, xyz, 4, 0)
    assert list(idx[2]) == [2, 2, 2, 2] and d[2, 0] == 0 and np.isinf(d[2, 1:]).all()
    assert list(idx[0]) == [0, 1, 0, 0] and np.isinf(d[0, 2:]).all()


def test_grid_respects_batch(oracle):
    xyz = np.array([[0, 0, 0], [0.1, 0, 0], [0.2, 0, 0]], np.float32)
    g = np.zeros((3, 3), np.int64)
    idx, _ = oracle.flashknn_query(g, np.array([0, 1, 0]), xyz, 3, 0)
    assert list(idx[0]) == [0, 2, 0]
    assert list(idx[1]) == [1, 1, 1]


def test_grid_num_down_parent_cells(oracle):
    # voxel 0 and voxel 3 share parent cell 0 at num_down=2; voxel 8 is two
    # parent cells away (parent 2) and is not a candidate of voxel 0.
    g = np.array([[0, 0, 0], [3, 0, 0], [8, 0, 0], [4, 0, 0]], np.int64)
    xyz = g.astype(np.float32)
    idx, d = oracle.flashknn_query(g, None, None, 4, 2)
    assert set(idx[0][np.isfinite(d[0])]) == {0, 1, 3}
    assert set(idx[2][np.isfinite(d[2])]) == {2, 3}
    assert np.array_equal(oracle.flashknn_query(g, None, xyz, 4, 2)[0], idx)


def test_selected_and_back_query_shapes(oracle):
    rng = np.random.default_rng(6)
    xyz = _cloud(rng, 400, 2.0)
    g = np.floor(xyz / np.float32(0.25)).astype(np.int64)
    qi = np.arange(0, 400, 7)
    si, sd = oracle.selected_query(xyz, g, qi, None, 8, 1)
    fi, fd = oracle.flashknn_query(g, None, xyz, 8, 1)
    assert np.array_equal(si, fi[qi]) and np.array_equal(sd, fd[qi])
    sel = np.arange(0, 400, 3)
    bi, bd = oracle.back_query(xyz, sel, 0.5, None, 4)
    assert bi.shape == (400, 4)
    # every point in sel finds itself (position in sel) at distance 0
    assert (bi[sel, 0] == np.arange(len(sel))).all() and (bd[sel, 0] == 0).all()


# ------------------------------------------------------- closed forms

def test_dist_backward_matches_finite_differences(oracle):
    rng = np.random.default_rng(7)
    q = rng.random((6, 3))
    p = rng.random((10, 3))
    idx = rng.integers(0, 10, (6, 4))
    g = rng.standard_normal((6, 4))

    def loss(qq, pp):
        d = ((qq[:, None, :] - pp[idx]) ** 2).sum(-1)
        return (d * g).sum()

    gq, gp = oracle.dist_backward(q, p, idx, g)
    h = 1e-6
    for arr, gr in ((q, gq), (p, gp)):
        num = np.zeros_like(arr)
        for i in range(arr.shape[0]):
            for a in range(3):
                old = arr[i, a]
                arr[i, a] = old + h
                lp = loss(q, p)
                arr[i, a] = old - h
                lm = loss(q, p)
                arr[i, a] = old
                num[i, a] = (lp - lm) / (2 * h)
        np.testing.assert_allclose(gr, num, rtol=1e-7, atol=1e-8)


def test_idw_weights_rows_sum_to_one(oracle):
    rng = np.random.default_rng(8)
    d2 = rng.random((5, 4))
    d2[0, 0] = 0.0
    w = oracle.idw_weights(d2, 1e-8)
    np.testing.assert_allclose(w.sum(1), 1.0)
    assert w[0, 0] > 0.999  # d2 = 0 dominates but stays finite


def test_point_candidates_vs_bruteforce(oracle):
    rng = np.random.default_rng(30)
    cell = rng.integers(-3, 4, (300, 3))
    batch = rng.integers(0, 2, 300)
    got = oracle.point_candidates(cell, batch)
    d = np.abs(cell[:, None, :] - cell[None, :, :]).max(-1)
    same = batch[:, None] == batch[None, :]
    assert np.array_equal(got, ((d <= 1) & same).sum(1))


def test_adaptive_selection_rule(oracle):
    # a dense blob and a sparse line: dense points descend to fine levels,
    # sparse ones stay coarse; every point's candidate set at its level is
    # at least min_f * k unless it sits at the coarsest level
    rng = np.random.default_rng(31)
    dense = rng.integers(0, 8, (400, 3))
    sparse = np.stack([np.arange(0, 600, 20), np.zeros(30, int), np.zeros(30, int)], 1)
    g = np.concatenate([dense, sparse]).astype(np.int64)
    idx, d, sel = oracle.adaptive_query(g, None, None, 8)
    top = int(g.max()).bit_length()
    assert sel[:400].mean() < sel[400:].mean()
    for s in np.unique(sel):
        q = np.where(sel == s)[0]
        cnt = oracle.point_candidates(g >> s, None)[q]
        assert (s == top) or (cnt >= 16).all()
    # rows are sorted and padded with the point itself
    assert (np.diff(d, axis=1) >= 0).all()


# ------------------------------------------------- voxel-first subsampling

def test_voxel_first_hand_vectors(oracle):
    pts = np.array([[0.01, 0, 0], [0.02, 0, 0], [0.5, 0, 0], [0.03, 0, 0]], np.float32)
    assert oracle.voxel_first(pts, 0.1).tolist() == [0, 2]
    # voxel coordinates are keyed by their low 16 bits (xyz2key, depth 16):
    # x = 0.5 and 65536.5 share voxel key 0; -0.5 (v = -1) and 65535.5 share 0xffff
    a = np.array([[0.5, 0, 0], [65536.5, 0, 0], [-0.5, 0, 0], [65535.5, 0, 0],
                  [0.5, 1.5, 0]], np.float32)
    assert oracle.voxel_first(a, 1.0).tolist() == [0, 2, 4]
    # clouds are voxelised separately
    assert oracle.voxel_first(a[:2], 1.0, offsets=[1, 2]).tolist() == [0, 1]


def test_voxel_first_true_quotient(oracle):
    """floor(x / size) with the correctly rounded fp32 quotient; the product
    with the fp32 reciprocal floors higher on this value."""
    x = np.array([[5.0399995, 0, 0], [5.04, 0, 0]], np.float32)
    assert np.floor(x[0, 0] / np.float32(0.06)) == 83
    assert np.floor(x[0, 0] * (np.float32(1) / np.float32(0.06))) == 84
    assert oracle.voxel_keys(x, 0.06)[0] != oracle.voxel_keys(x, 0.06)[1]
    assert oracle.voxel_first(x, 0.06).tolist() == [0, 1]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_voxel_first_vs_python_loop(oracle, seed):
    rng = np.random.default_rng(seed)
    pts = (rng.random((3000, 3), dtype=np.float32) * 4 - 2).astype(np.float32)
    off = np.array([700, 700, 2100, 3000], np.int32)
    for o in (None, off):
        got = oracle.voxel_first(pts, 0.3, o)
        assert np.array_equal(got, oracle.voxel_first_python(pts, 0.3, o))
        # invariants: one point per voxel key; every key represented by an
        # index no larger than any of its points
        key = oracle.voxel_keys(pts, 0.3, o)
        assert len(np.unique(key[got])) == len(got) == len(np.unique(key))
        first = {}
        for i, k in enumerate(key.tolist()):
            first.setdefault(k, i)
        assert sorted(first.values()) == got.tolist()


# ------------------------------------ adaptive level rule, hand-computed
# AdaptiveNeighborhoodFlashKNN's level rule
# (FlashKNN/functions/adaptive_octree.py:327-350): start at the coarsest
# level; a point is active there if its 27-node count is > max_f * k; at
# each finer level an active point takes the level if its count there is
# >= min_f * k, and stays active only if that count is also > max_f * k.
#
# Hand case, all points on the x axis (y = z = 0), so a node's 27-node count
# is the number of same-batch points whose x >> s differs by at most 1:
#   batch 0: A = 12 points at x 0, B = 3 at x 1, G = 1 at x 3,
#            D = 1 at x 12, E = 1 at x 15          (max x 15: levels 0..4)
#   batch 1: H = 8 points at x 7 (shifted to 0 per batch)
# counts per level s (hand-enumerated):
#        s=4  s=3  s=2  s=1  s=0
#   A    18   18   16   16   15     (s=1: A,B in node 0, G in node 1)
#   B    18   18   16   16   15     (s=0: A at 0, B at 1, G at 3 -> not G)
#   G    18   18   16   16    1     (s=0: nodes 2..4 hold only G)
#   D    18   18    2    2    1     (s=2: D,E share node 3; s=1: nodes 6,7)
#   E    18   18    2    2    1
#   H     8    8    8    8    8
_ADAPT_X = [0] * 12 + [1] * 3 + [3, 12, 15] + [7] * 8
_ADAPT_B = [0] * 18 + [1] * 8
_ADAPT_COUNTS = {  # level -> per group (A, B, G, D, E, H)
    4: (18, 18, 18, 18, 18, 8), 3: (18, 18, 18, 18, 18, 8),
    2: (16, 16, 16, 2, 2, 8), 1: (16, 16, 16, 2, 2, 8), 0: (15, 15, 1, 1, 1, 8)}
_ADAPT_GROUPS = [12, 3, 1, 1, 1, 8]


def _adapt_case():
    g = np.zeros((len(_ADAPT_X), 3), np.int64)
    g[:, 0] = _ADAPT_X
    return g, np.array(_ADAPT_B, np.int64)


def _per_group(vals):
    return np.repeat(np.array(vals, np.int64), _ADAPT_GROUPS)


def test_adaptive_hand_counts(oracle):
    g, b = _adapt_case()
    gn = g - np.array([[7, 0, 0]]) * (b[:, None] == 1)  # per-batch shift
    for s, want in _ADAPT_COUNTS.items():
        assert oracle.point_candidates(gn >> s, b).tolist() == _per_group(want).tolist(), s


@pytest.mark.parametrize("k,expect", [
    # k = 1 (thresholds 2 and 8):
    #   A, B descend to 0 (15 >= 2 at every level, > 8 down to s = 1);
    #   G: 16 > 8 at s = 1 but 1 < 2 at s = 0 -> falls back to 1;
    #   D, E: 2 at s = 2 is >= 2 but not > 8 -> stop at 2 (s = 1 also has 2:
    #         a rule that kept descending would pick 1);
    #   H: 8 at the coarsest level is not > 8 -> stays at 4.
    (1, (0, 0, 1, 2, 2, 4)),
    # k = 2 (thresholds 4 and 16): A, B, G take s = 2 (16 >= 4) and stop
    #   there (16 is not > 16); D, E: 2 < 4 at s = 2 -> stay at 3; H stays at 4.
    (2, (2, 2, 2, 3, 3, 4)),
])
def test_adaptive_hand_levels(oracle, k, expect):
    g, b = _adapt_case()
    _, d, sel = oracle.adaptive_query(g, b, None, k)
    assert sel.tolist() == _per_group(expect).tolist()
    assert np.isfinite(d[:, 0]).all()  # every point finds itself at least


# ----------------------------------- back_query cells truncate toward zero
# back_query takes cells (xyz / query_grid_size).long()
# (FlashKNN/functions/FlashKnnWrapper.py:480): truncation, so (-size, size)
# is one double-width cell around 0.  Hand case on the x axis, size 0.1:
#   p0 x =  0.15 -> cell  1     (support, position 0 in sel)
#   p1 x = -0.05 -> cell  0     (floor would give -1)
#   p2 x = -0.25 -> cell -2     (support, position 1; floor: -3)
#   p3 x = -0.12 -> cell -1     (floor: -2)
# K = 1 neighbours (position in sel; candidates within one cell):
#   p0 -> p0;  p1 -> p0 (cells 0 and 1);  p2 -> p2;  p3 -> p2 (cells -1, -2).
# With floor cells p1 (cell -1) would see neither p0 (cell 1) nor p2
# (cell -3) and get the padding -1.
def test_back_query_truncates_negative_cells(oracle):
    xyz = np.zeros((4, 3), np.float32)
    xyz[:, 0] = [0.15, -0.05, -0.25, -0.12]
    sel = np.array([0, 2], np.int64)
    i1, d1 = oracle.back_query(xyz, sel, 0.1, None, 1)
    assert i1[:, 0].tolist() == [0, 0, 1, 1]
    i2, d2 = oracle.back_query(xyz, sel, 0.1, None, 2)
    assert i2.tolist() == [[0, -1], [0, -1], [1, -1], [1, -1]]
    assert np.isinf(d2[:, 1]).all() and np.isfinite(d2[:, 0]).all()
    assert d2[0, 0] == 0 and d2[2, 0] == 0
