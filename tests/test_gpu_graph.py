"""The differentiable hierarchy step (exact kNN between consecutive levels,
IDW interpolation, backward to queries, supports and features) captured
into one graph with torch.cuda.graph (needs an MI355X).

Every replay must reproduce the eager step bit for bit (indices and every
gradient), with eager steps run on the same kNN engines between replays;
the replayed indices are also checked against the oracle's brute force on
a fixed sample of queries per level.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

fk = pytest.importorskip("flashknn_exprepo_amd")
from flashknn_exprepo_amd import synthetic  # noqa: E402


def _dev():
    return torch.device("cuda", 0)


def _make_step(stages, feats, wts, offs, knns, k=16):
    def step():
        leaves, idxs, outs, grads = [], [], [], []
        for l in range(len(stages) - 1):
            q = stages[l].clone().requires_grad_(True)
            p = stages[l + 1].clone().requires_grad_(True)
            f = feats[l + 1].clone().requires_grad_(True)
            idx, _ = knns[l](p.detach(), offs[l + 1], k, q.detach(), offs[l])
            outs.append(fk.idw_interpolate(f, q, p, idx, eps=1e-8))
            grads.append(wts[l])
            leaves += [q, p, f]
            idxs.append(idx)
        torch.autograd.backward(outs, grads)
        return idxs, outs, leaves
    return step


def _snapshot(res):
    idxs, outs, leaves = res
    return ([i.clone() for i in idxs], [o.detach().clone() for o in outs],
            [t.grad.clone() for t in leaves])


def _same(a, b, what):
    for x, y in zip(a, b):
        if x.dtype == torch.float32:
            assert torch.equal(x.view(torch.int32), y.view(torch.int32)), what
        else:
            assert torch.equal(x, y), what


def test_hierarchy_step_graph_replays_match_eager(oracle):
    dev = _dev()
    stages_np = synthetic.hierarchy(seed=51)
    stages = [torch.from_numpy(s).to(dev) for s in stages_np]
    g = torch.Generator(device="cpu").manual_seed(5)
    feats = [torch.randn(s.shape[0], 64, generator=g).to(dev) for s in stages]
    wts = [torch.randn(s.shape[0], 64, generator=g).to(dev) for s in stages]
    offs = [torch.tensor([s.shape[0]], dtype=torch.int32, device=dev) for s in stages]
    knns = [fk.StaticKNN() for _ in range(len(stages) - 1)]
    step = _make_step(stages, feats, wts, offs, knns)

    print("graph: eager step (freezes cell sizes)", flush=True)
    first = _snapshot(step())
    print("graph: eager static step", flush=True)
    ref = _snapshot(step())
    _same(ref[0], first[0], "static indices differ from the first call")
    _same(ref[2], first[2], "static gradients differ from the first call")

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    print("graph: capture", flush=True)
    with torch.cuda.graph(graph):
        captured = step()
    for r in range(3):
        print(f"graph: replay {r}", flush=True)
        graph.replay()
        torch.cuda.synchronize()
        got = _snapshot(captured)
        _same(got[0], ref[0], f"replay {r}: indices")
        _same(got[1], ref[1], f"replay {r}: outputs")
        _same(got[2], ref[2], f"replay {r}: gradients")
        # eager work on the same engines between replays
        _same(_snapshot(step())[0], ref[0], f"eager step after replay {r}")
    for kn in knns:
        kn.status()

    rng = np.random.default_rng(0)
    for l in range(len(stages_np) - 1):
        q, p = stages_np[l], stages_np[l + 1]
        rows = rng.choice(len(q), min(2000, len(q)), replace=False)
        bi, _ = oracle.brute_knn(p, [len(p)], q[rows], [len(rows)], 16)
        assert np.array_equal(ref[0][l].cpu().numpy()[rows], bi), f"level {l} vs oracle"
