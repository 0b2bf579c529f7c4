"""Benchmark of the FlashKNN grid-kNN path on one MI355X.

Headline (value): exact kNN query throughput, M queries/s, on a synthetic
S3DIS-like room of ~100k fp32 points, queries = all points, K = 16, idx
bit-exact against brute force (the first pass scans the 27-cell block, the
queries it cannot certify are finished by shell expansion; both kernels are
inside the timed step).  Inputs and the built grid are resident in HBM; the
grid construction is timed separately (``extra.build_ms``).

One step = one full query of all points.  W warmup steps, then K timed
steps bracketed by torch.cuda.synchronize() (one GPU, no barrier needed).

Also reported in ``extra``: K=32, the LiDAR sweep, the uniform 8192 case,
the reference's 27-cell FlashKNN.query semantics, the 16-cloud batched
workload (graph replay vs 16 per-cloud launches), and the differentiable
hierarchy fwd+bwd.  ``roofline`` is for the dominant kernel (the first
pass, k_query<16>), timed alone with HIP events on its stream.
``cpu_baseline`` times the reference's CPU brute-force path (torch.cdist +
topk in query blocks of 10,000, as FlashKNN/functions/utils.py:5-82) on a
bounded query sample on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("M queries/s at 100k pts K=16 (bit-exact idx); batched vs per-cloud; "
          "fwd+bwd ms")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def events_ms(fn, reps, stream=None):
    """Median ms of fn() measured with HIP events on the current stream."""
    import torch
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
    for i in range(reps):
        ev[2 * i].record()
        fn()
        ev[2 * i + 1].record()
    torch.cuda.synchronize()
    t = sorted(ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(reps))
    return t[reps // 2], sum(t) / reps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run 5 launches of a plain fp32 copy kernel over this many "
                         "MiB (the known-size reference for hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "only one GPU is measured: the batched kNN path "
                          "shards by cloud (each GPU takes whole clouds, no exchange), "
                          "but no multi-GPU figure is produced", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import flashknn_exprepo_amd as fk
    from flashknn_exprepo_amd import _lib as F
    from flashknn_exprepo_amd import synthetic

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    fk.load()
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731

    # ---------------------------------------------------------- headline
    log("generating room (~100k points)")
    room = synthetic.room(seed=48)
    n = len(room)
    xyz = T(room)
    off = T(np.array([n], np.int32))
    K = 16
    eng = fk.ExactKNN()
    cell = eng.build(xyz, off, K)          # density first guess + feedback
    inv = (1.0 / cell.double()).float()
    g = eng.grid
    log("build")
    build = lambda: g.build_xyz(xyz, off, inv)  # noqa: E731
    build()
    build_ms, _ = events_ms(build, 10)
    # int32 indices (the pointops layout; 4 bytes per neighbour index as in
    # the algorithmic byte count) and fp32 squared distances
    out_i = torch.empty((n, K), dtype=torch.int32, device=dev)
    out_d = torch.empty((n, K), dtype=torch.float32, device=dev)

    def step():
        g.query(K, mode=F.FK_MODE_EXACT, pad_mode=F.FK_PAD_MINUS1, out_idx=out_i,
                out_dist=out_d, **g.hints())

    log(f"headline: n={n} K={K} warmup={args.warmup} steps={args.steps}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3) / 1e6
    stats = g.query_stats()

    # bit-exactness of the measured output on a sample (oracle as checker)
    from oracle import oracle as orc
    sample = np.random.default_rng(0).choice(n, 2000, replace=False)
    ri, rd = orc.brute_knn(room, [n], room[sample], [2000], K)
    exact_ok = bool(np.array_equal(out_i.cpu().numpy()[sample], ri))

    # dominant kernel alone (first pass), HIP events on the current stream
    # (the stream g.query launches on) around K back-to-back launches: each
    # first pass clears the counter set the next one uses, so no reset
    # launch runs in between, and the average includes the launch gaps
    def pass1():
        g.query(K, mode=F.FK_MODE_EXACT, pad_mode=F.FK_PAD_MINUS1, out_idx=out_i,
                out_dist=out_d, passes=2, **g.hints())

    for _ in range(args.warmup):
        pass1()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        pass1()
    e1.record()
    torch.cuda.synchronize()
    p1_ms = e0.elapsed_time(e1) / args.steps
    ncells = g.info()["cells"]
    bytes_q = 12 * n + 12 * n + 8 * K * n + 4 * (ncells + 1)
    achieved = bytes_q / (p1_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "query_room100k_k16_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": traffic, "kernel": "k_query_planned<16> (first pass, balanced plan)",
                "kernel_ms": round(p1_ms, 5), "algorithmic_bytes": bytes_q,
                "timing": f"HIP events around {args.steps} back-to-back first passes"}

    if args.copy_mb > 0:
        # torch.neg: one elementwise kernel reading and writing copy_mb MiB
        a = torch.randn(args.copy_mb * (1 << 18), device=dev)
        b = torch.empty_like(a)
        for _ in range(5):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        del a, b

    extra = {"build_ms": round(build_ms, 4), "cells": ncells,
             "certified_first_pass": stats["certified_first_pass"],
             "expanded": stats["expanded"], "bitexact_sample_2000": exact_ok,
             "cell_side_m": round(float(cell[0]), 5), "prune": bool(g.prune_hint),
             "one_lane": bool(g.one_lane_hint), "candidate_batch": g.batch_hint}

    if not args.quick:
        def qps(grid, k, mode, reps=20, **kw):
            oi = torch.empty((grid.num_points, k), dtype=torch.int64, device=dev)
            od = torch.empty((grid.num_points, k), dtype=torch.float32, device=dev)
            f = lambda: grid.query(k, mode=mode, out_idx=oi, out_dist=od,  # noqa: E731
                                   **grid.hints(), **kw)
            f()
            med, _ = events_ms(f, reps)
            return grid.num_points / (med * 1e-3) / 1e6, med

        # K = 32 on the same room
        e32 = fk.ExactKNN()
        e32.build(xyz, off, 32)
        v, ms = qps(e32.grid, 32, F.FK_MODE_EXACT, pad_mode=F.FK_PAD_MINUS1)
        extra["room_k32_Mqps"], extra["room_k32_ms"] = round(v, 2), round(ms, 4)

        # reference semantics: FlashKNN.query, voxel 4 cm, num_down 2 (27 cells)
        grid_c = torch.floor(xyz / 0.04).long()
        gc = fk.Grid()
        gc.build_cells(grid_c, None, xyz, num_down=2)
        v, ms = qps(gc, K, F.FK_MODE_CELLS)
        extra["room_27cell_alpha4_Mqps"], extra["room_27cell_alpha4_ms"] = round(v, 2), round(ms, 4)
        extra["room_27cell_mean_candidates"] = round(gc.candidate_count(1) / n, 2)

        # orientation at the reference's published size (~250k points): the
        # 27-cell alpha-4 query (FlashKNN.query semantics) at K=16 and K=32,
        # and the exact query at K=16
        log("room 250k")
        r250 = synthetic.room(seed=48, target=250000)
        x250 = T(r250)
        g250 = fk.Grid()
        g250.build_cells(torch.floor(x250 / 0.04).long(), None, x250, num_down=2)
        extra["room250k_n"] = len(r250)
        for kk in (16, 32):
            v, ms = qps(g250, kk, F.FK_MODE_CELLS)
            extra[f"room250k_27cell_alpha4_k{kk}_Mqps"] = round(v, 2)
            extra[f"room250k_27cell_alpha4_k{kk}_ms"] = round(ms, 4)
        e250 = fk.ExactKNN()
        e250.build(x250, T(np.array([len(r250)], np.int32)), K)
        v, ms = qps(e250.grid, K, F.FK_MODE_EXACT, pad_mode=F.FK_PAD_MINUS1)
        extra["room250k_exact_k16_Mqps"], extra["room250k_exact_k16_ms"] = round(v, 2), round(ms, 4)
        del g250, e250

        # LiDAR sweep
        log("lidar")
        lid = synthetic.lidar(seed=49)
        lx = T(lid)
        lo = T(np.array([len(lid)], np.int32))
        el = fk.ExactKNN()
        el.build(lx, lo, K)
        v, ms = qps(el.grid, K, F.FK_MODE_EXACT, pad_mode=F.FK_PAD_MINUS1)
        extra["lidar_n"] = len(lid)
        extra["lidar_k16_Mqps"], extra["lidar_k16_ms"] = round(v, 2), round(ms, 4)
        extra["lidar_prune"] = bool(el.grid.prune_hint)
        # the LiDAR config's own semantics: the adaptive neighbourhood
        # (per-point octree level by the 2K/8K candidate rule) on 5 cm voxel
        # coordinates; whole call timed (level builds + selection + query),
        # set recall against the exact kNN
        exact_idx, _ = el(lx, lo, K)
        gl = torch.floor(lx / 0.05).long()
        lb = torch.zeros(len(lid), dtype=torch.long, device=dev)
        an = fk.AdaptiveNeighborhoodFlashKNN(num_nbr=K)
        ad_idx = an.query(gl, lb, lx)
        ad_med, _ = events_ms(lambda: an.query(gl, lb, lx), 10)
        extra["lidar_adaptive_k16_Mqps"] = round(len(lid) / (ad_med * 1e-3) / 1e6, 2)
        extra["lidar_adaptive_k16_ms"] = round(ad_med, 4)
        extra["lidar_adaptive_recall"] = round(float(fk.cal_recall(ad_idx, exact_idx).mean()), 5)

        # uniform 8192
        uni = synthetic.uniform(8192, 47)
        ux, uo = T(uni), T(np.array([8192], np.int32))
        eu = fk.ExactKNN()
        eu.build(ux, uo, K)
        v, ms = qps(eu.grid, K, F.FK_MODE_EXACT, pad_mode=F.FK_PAD_MINUS1)
        extra["uniform8192_k16_Mqps"], extra["uniform8192_k16_ms"] = round(v, 2), round(ms, 4)

        # batched 16 uneven clouds: graph replay vs per-cloud launches
        log("batched")
        clouds, ks = synthetic.batch(seed=50)
        bx, bo = synthetic.pack(clouds)
        bk = fk.BatchedKNN(T(bx), T(bo), ks, exact=True)
        bk.run()
        bk.capture()
        rep_med, _ = events_ms(bk.replay, 20)
        b_idx = bk.out_idx.clone()
        engines = []
        start = 0
        for c, (cl, k) in enumerate(zip(clouds, ks)):
            e = fk.ExactKNN()
            cx = T(cl)
            co = T(np.array([len(cl)], np.int32))
            e.build(cx, co, k, cell=bk.cell[c:c + 1], prune=bk.grid.prune_hint)
            oi = torch.empty((len(cl), k), dtype=torch.int64, device=dev)
            od = torch.empty((len(cl), k), dtype=torch.float32, device=dev)
            engines.append((e.grid, k, oi, od, start))
            start += len(cl)

        def per_cloud():
            for gr, k, oi, od, _ in engines:
                gr.query(k, mode=F.FK_MODE_EXACT, pad_mode=F.FK_PAD_SELF, out_idx=oi,
                         out_dist=od, **gr.hints())

        pc_med, _ = events_ms(per_cloud, 20)
        same = True
        for c, (gr, k, oi, od, st) in enumerate(engines):
            bi, _ = bk.per_cloud(c)
            same &= bool(torch.equal(bi - st, oi))
        nb = len(bx)
        extra.update({"batched_points": nb, "batched_graph_ms": round(rep_med, 4),
                      "batched_Mqps": round(nb / (rep_med * 1e-3) / 1e6, 2),
                      "per_cloud_ms": round(pc_med, 4),
                      "per_cloud_Mqps": round(nb / (pc_med * 1e-3) / 1e6, 2),
                      "batched_speedup_vs_per_cloud": round(pc_med / rep_med, 3),
                      "batched_idx_bitexact_vs_per_cloud": same,
                      "batched_k": ks})
        del b_idx

        # differentiable hierarchy fwd+bwd
        log("hierarchy")
        stages = [T(s) for s in synthetic.hierarchy(seed=51)]
        feats = [torch.randn(s.shape[0], 64, device=dev) for s in stages]
        wts = [torch.randn(s.shape[0], 64, device=dev) for s in stages]

        offs = [torch.tensor([s.shape[0]], dtype=torch.int32, device=dev) for s in stages]
        knns = [fk.StaticKNN() for _ in range(len(stages) - 1)]

        def fwd_bwd():
            # loss = sum_l <out_l, wts_l>: its gradient w.r.t. out_l is wts_l
            outs, leaves = [], []
            for l in range(len(stages) - 1):
                q = stages[l].clone().requires_grad_(True)
                p = stages[l + 1].clone().requires_grad_(True)
                f = feats[l + 1].clone().requires_grad_(True)
                leaves += [q, p, f]
                idx, _ = knns[l](p.detach(), offs[l + 1], 16, q.detach(), offs[l])
                outs.append(fk.idw_interpolate(f, q, p, idx, eps=1e-8))
            torch.autograd.backward(outs, wts[:len(outs)])
            return leaves

        fwd_bwd()  # freezes the per-level cell sizes
        h_eager, _ = events_ms(fwd_bwd, 10)
        # the steady-state step captured once into a graph and replayed
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            fwd_bwd()
        torch.cuda.current_stream().wait_stream(side)
        hgraph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(hgraph):
            cap_leaves = fwd_bwd()
        ref_leaves = fwd_bwd()
        hgraph.replay()
        torch.cuda.synchronize()
        same_h = all(torch.equal(a.grad, b.grad) for a, b in zip(cap_leaves, ref_leaves))
        h_med, _ = events_ms(hgraph.replay, 20)
        for kn in knns:
            kn.status()
        extra["hierarchy_sizes"] = [int(s.shape[0]) for s in stages]
        extra["hierarchy_fwd_bwd_ms"] = round(h_med, 4)
        extra["hierarchy_fwd_bwd_eager_ms"] = round(h_eager, 4)
        extra["hierarchy_graph_grads_equal_eager"] = bool(same_h)

    # ------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cores = os.cpu_count()
        torch.set_num_threads(cores)
        s = torch.from_numpy(room)
        qn = 8192
        qs = torch.from_numpy(room[np.random.default_rng(1).choice(n, qn, replace=False)])

        def cpu_run():
            for b0 in range(0, qn, 10000):
                d = torch.cdist(qs[b0:b0 + 10000], s)
                torch.topk(d, K, dim=1, largest=False)

        for _ in range(3):  # warm-ups
            cpu_run()
        reps = []
        for _ in range(10):
            t0 = time.perf_counter()
            cpu_run()
            reps.append(time.perf_counter() - t0)
        tc = sorted(reps)[len(reps) // 2]
        model = "unknown"
        try:
            with open("/proc/cpuinfo") as f:
                model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        cpu = {"value": round(qn / tc / 1e6, 5), "unit": "M queries/s",
               "cores": torch.get_num_threads(), "kind": "port",
               "cpu_model": model, "os_cpu_count": cores,
               "sample": f"{qn} random queries of the {n}-point room against all "
                         f"{n} points, torch.cdist+topk in blocks of 10000, "
                         f"3 warm-ups + median of 10"}

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "M queries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"synthetic indoor room, {n} fp32 points (4 cm voxel "
                               "subsampled), queries = all points, K=16, exact kNN "
                               "(bit-exact idx; int32 idx + fp32 squared distance "
                               "written), single cloud",
                   "points": n, "k": K, "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
