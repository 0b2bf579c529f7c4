"""The reference's brute-force and recall helpers on device tensors
(FlashKNN/functions/utils.py:5-104), as FlashKNN/test.py calls them: on
the GPU, against the oracle's brute force and the HIP exact kNN.

* vanilla_Knn keeps the matmul expansion ||q||^2 + ||s||^2 - 2 q.s, whose
  fp32 rounding may swap near ties; its rows are compared as sets with
  the fp64 brute force (>= 99 % of rows equal, distances within 1e-3
  relative / 1e-4 absolute), and its CPU and GPU runs must agree on the
  same inputs up to that rounding.
* cal_recall is integer work: on device tensors it must equal the
  hand-computed multiset recalls and its own CPU result bit for bit.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

fk = pytest.importorskip("flashknn_exprepo_amd")
from flashknn_exprepo_amd.functions import cal_recall, vanilla_Knn  # noqa: E402


def _dev():
    return torch.device("cuda", 0)


def test_cal_recall_device_hand_cases():
    gt = torch.tensor([[1, 2, 3, 4], [5, 5, 6, 7], [1, 2, 3, 4], [0, 0, 0, 0]])
    pr = torch.tensor([[4, 3, 2, 1], [5, 6, 6, 9], [9, 9, 9, 1], [0, 0, 1, 2]])
    r = cal_recall(pr.to(_dev()), gt.to(_dev()))
    assert r.is_cuda
    # row 1: 5 once (gt has two), 6 once (pred has two) -> 2/4; row 3: 0 twice
    assert r.cpu().tolist() == [1.0, 0.5, 0.25, 0.5]


def test_cal_recall_device_equals_cpu():
    rng = np.random.default_rng(3)
    gt = torch.from_numpy(rng.integers(0, 40, (5000, 16)))
    pr = torch.from_numpy(rng.integers(0, 40, (5000, 16)))
    a = cal_recall(pr, gt)
    b = cal_recall(pr.to(_dev()), gt.to(_dev())).cpu()
    assert torch.equal(a, b)


@pytest.mark.parametrize("k", [8, 16])
def test_vanilla_knn_device_vs_oracle(oracle, k):
    rng = np.random.default_rng(10 + k)
    s = rng.random((6000, 3), dtype=np.float32)
    q = rng.random((1500, 3), dtype=np.float32)
    d, i = vanilla_Knn(torch.from_numpy(q).to(_dev()), torch.from_numpy(s).to(_dev()), k)
    assert d.is_cuda and i.is_cuda and d.shape == (1500, k) and i.dtype == torch.int64
    ri, rd = oracle.brute_knn(s, [6000], q, [1500], k, f64=True)
    gi = i.cpu().numpy()
    same = (np.sort(gi, 1) == np.sort(ri, 1)).all(1).mean()
    assert same > 0.99, same
    np.testing.assert_allclose(d.cpu().numpy(), np.sqrt(rd), rtol=1e-3, atol=1e-4)
    # the same helper on the CPU gives the same neighbours up to that rounding
    dc, ic = vanilla_Knn(torch.from_numpy(q), torch.from_numpy(s), k)
    same_cpu = (np.sort(ic.numpy(), 1) == np.sort(gi, 1)).all(1).mean()
    assert same_cpu > 0.99, same_cpu


def test_recall_of_hip_exact_knn_is_one(oracle):
    """FlashKNN/test.py's use: the recall of the HIP exact kNN against
    vanilla_Knn on the device, and against the oracle's fp32 brute force
    (bit-exact rows, so recall exactly 1)."""
    rng = np.random.default_rng(21)
    xyz = rng.random((8192, 3), dtype=np.float32)
    x = torch.from_numpy(xyz).to(_dev())
    off = torch.tensor([8192], dtype=torch.int32, device=_dev())
    idx, _ = fk.knn(x, off, 16)
    bi, _ = oracle.brute_knn(xyz, [8192], xyz, [8192], 16)
    r = cal_recall(idx, torch.from_numpy(bi).to(_dev()))
    assert bool((r == 1.0).all())
    _, vi = vanilla_Knn(x, x, 16)
    rv = cal_recall(idx, vi)
    assert float(rv.mean()) > 0.999
