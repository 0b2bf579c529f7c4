"""fk_voxel_first_noalloc (the library's static per-GPU workspace, no
allocation) against fk_voxel_first with a caller workspace, through the C
ABI (needs an MI355X).  Calls in a row of decreasing size share the one
static array, so each must be exact after a larger one."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

fk = pytest.importorskip("flashknn_exprepo_amd")
from flashknn_exprepo_amd import _lib as F  # noqa: E402
from flashknn_exprepo_amd import synthetic  # noqa: E402


def _dev():
    return torch.device("cuda", 0)


def _voxel(L, xyz, off, B, size, noalloc):
    n = xyz.shape[0]
    out = torch.full((max(n, 1),), -7, dtype=torch.int64, device=_dev())
    cnt = torch.full((1,), -7, dtype=torch.int64, device=_dev())
    if noalloc:
        st = L.fk_voxel_first_noalloc(F.ptr(xyz), F.ptr(off), B, n, ctypes.c_float(size),
                                      F.ptr(out), F.ptr(cnt), F.stream_ptr())
        F.check(st, "fk_voxel_first_noalloc")
    else:
        ws = torch.empty(L.fk_voxel_workspace_bytes(n), dtype=torch.uint8, device=_dev())
        st = L.fk_voxel_first(F.ptr(xyz), F.ptr(off), B, n, ctypes.c_float(size), F.ptr(ws),
                              F.ptr(out), F.ptr(cnt), F.stream_ptr())
        F.check(st, "fk_voxel_first")
    torch.cuda.synchronize()
    c = int(cnt.item())
    assert c > 0, c
    return out[:c].cpu().numpy()


def _cases():
    rng = np.random.default_rng(61)
    nmax = F.FK_VOXEL_NOALLOC_MAX_POINTS
    uni = (rng.random((nmax, 3), dtype=np.float32) * 10).astype(np.float32)
    room = synthetic.room(seed=48)
    three = np.concatenate([room, room[::2] + np.float32(0.013), room[::3]])
    off3 = np.cumsum([len(room), len(room[::2]), len(room[::3])]).astype(np.int32)
    small = room[:20000]
    # largest first: later calls run on the array the larger one left behind
    return [("uniform at the maximum size", uni, np.array([nmax], np.int32), 1, 0.05),
            ("three room clouds", three, off3, 3, 0.05),
            ("one room cloud, no offsets", small, None, 1, 0.1)]


def test_noalloc_matches_workspace_entry_point(oracle):
    L = F.load()
    for tag, x, off, B, size in _cases():
        print(f"case: {tag} n={len(x)} clouds={B} size={size}", flush=True)
        xyz = torch.from_numpy(x).to(_dev())
        o = None if off is None else torch.from_numpy(off).to(_dev())
        ref = _voxel(L, xyz, o, B, size, noalloc=False)
        got = _voxel(L, xyz, o, B, size, noalloc=True)
        assert np.array_equal(got, ref), tag
        if len(x) <= 400_000:
            assert np.array_equal(got, oracle.voxel_first(x, size, off)), tag


def test_noalloc_refuses_above_maximum():
    L = F.load()
    n = F.FK_VOXEL_NOALLOC_MAX_POINTS + 1
    xyz = torch.zeros((n, 3), dtype=torch.float32, device=_dev())
    out = torch.full((n,), -7, dtype=torch.int64, device=_dev())
    cnt = torch.full((1,), -7, dtype=torch.int64, device=_dev())
    st = L.fk_voxel_first_noalloc(F.ptr(xyz), None, 1, n, ctypes.c_float(0.1), F.ptr(out),
                                  F.ptr(cnt), F.stream_ptr())
    assert st == F.FK_ERR_CAPACITY
    torch.cuda.synchronize()
    assert int(cnt.item()) == -7 and bool((out == -7).all())
