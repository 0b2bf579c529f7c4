"""Seeded synthetic point clouds for the benchmark and test workloads.

* ``uniform(n, seed)``: i.i.d. U[0,1]^3.
* ``room(seed, target)``: a 6 x 8 x 3 m shell (floor, ceiling, four walls)
  plus 20 axis-aligned boxes with sides U(0.3, 2) m, points uniform on all
  surfaces by area, Gaussian noise sigma = 5 mm, voxel-first subsampled at
  4 cm; the raw count is calibrated so about ``target`` points remain.
* ``lidar(seed)``: 64 beams (elevation -25..+3 deg) x 2048 azimuth steps
  cast against a ground plane z = -1.7 m and 200 vertical cylinders and
  boxes, range <= 50 m, N(0, 2 cm) range noise, voxel-first subsampled at
  5 cm; density falls off as 1/r^2.
* ``batch(seed)``: 16 clouds, sizes log-uniform in [2k, 120k], alternating
  room / LiDAR generators seeded seed + i, per-cloud k in {8, 16, 32}.
* ``hierarchy(seed)``: a room of ~100k points subsampled at 12.5/25/50 cm
  (~25k, ~6k, ~1.5k points).

Voxel-first subsampling keeps, per voxel floor(xyz / size), the point with
the smallest index (the deterministic form of the reference's
``_voxel_first_cuda``, Networks/hierarchy.py:23-36).
"""
from __future__ import annotations

import numpy as np


def uniform(n=8192, seed=47):
    rng = np.random.default_rng(seed)
    return rng.random((n, 3), dtype=np.float32)


def voxel_first(xyz: np.ndarray, size: float) -> np.ndarray:
    """Indices (ascending) of the first point of every occupied voxel."""
    g = np.floor(xyz.astype(np.float64) / size).astype(np.int64)
    g -= g.min(0)
    key = (g[:, 0] * (g[:, 1].max() + 1) + g[:, 1]) * (g[:, 2].max() + 1) + g[:, 2]
    _, first = np.unique(key, return_index=True)
    return np.sort(first)


def _room_raw(rng, n_raw):
    W, L, H = 6.0, 8.0, 3.0
    quads = [  # origin, edge u, edge v
        ((0, 0, 0), (W, 0, 0), (0, L, 0)), ((0, 0, H), (W, 0, 0), (0, L, 0)),
        ((0, 0, 0), (W, 0, 0), (0, 0, H)), ((0, L, 0), (W, 0, 0), (0, 0, H)),
        ((0, 0, 0), (0, L, 0), (0, 0, H)), ((W, 0, 0), (0, L, 0), (0, 0, H)),
    ]
    for _ in range(20):
        s = rng.uniform(0.3, 2.0, 3)
        s[2] = min(s[2], H)
        o = np.array([rng.uniform(0, W - s[0]), rng.uniform(0, L - s[1]), 0.0])
        x, y, z = s
        quads += [
            (tuple(o), (x, 0, 0), (0, y, 0)), (tuple(o + [0, 0, z]), (x, 0, 0), (0, y, 0)),
            (tuple(o), (x, 0, 0), (0, 0, z)), (tuple(o + [0, y, 0]), (x, 0, 0), (0, 0, z)),
            (tuple(o), (0, y, 0), (0, 0, z)), (tuple(o + [x, 0, 0]), (0, y, 0), (0, 0, z)),
        ]
    O = np.array([q[0] for q in quads], np.float64)
    U = np.array([q[1] for q in quads], np.float64)
    V = np.array([q[2] for q in quads], np.float64)
    area = np.linalg.norm(np.cross(U, V), axis=1)
    pick = rng.choice(len(quads), n_raw, p=area / area.sum())
    a, b = rng.random(n_raw), rng.random(n_raw)
    pts = O[pick] + a[:, None] * U[pick] + b[:, None] * V[pick]
    pts += rng.normal(0, 0.005, pts.shape)
    return pts


def room(seed=48, target=100_000, voxel=0.04):
    """Synthetic indoor room, ~target points after voxel subsampling."""
    n_raw = int(target * 1.35)
    for _ in range(6):
        rng = np.random.default_rng(seed)
        raw = _room_raw(rng, n_raw)
        keep = voxel_first(raw, voxel)
        if abs(len(keep) - target) <= 0.02 * target:
            break
        # saturating map raw -> kept; adjust the raw count proportionally
        n_raw = max(1000, int(n_raw * (target / max(len(keep), 1)) ** 1.6))
    return raw[keep].astype(np.float32)


def lidar(seed=49, voxel=0.05, beams=64, steps=2048, rmax=50.0):
    rng = np.random.default_rng(seed)
    el = np.deg2rad(np.linspace(-25.0, 3.0, beams))
    az = np.linspace(0, 2 * np.pi, steps, endpoint=False)
    E, A = np.meshgrid(el, az, indexing="ij")
    d = np.stack([np.cos(E) * np.cos(A), np.cos(E) * np.sin(A), np.sin(E)], -1).reshape(-1, 3)
    sensor_z = 0.0
    t = np.full(len(d), np.inf)
    # ground plane z = -1.7
    down = d[:, 2] < -1e-9
    t[down] = (-1.7 - sensor_z) / d[down, 2]
    # 200 objects: vertical cylinders and axis-aligned boxes
    n_obj = 200
    ctr = rng.uniform(-45, 45, (n_obj, 2))
    far = np.linalg.norm(ctr, axis=1) > 3.0
    ctr = ctr[far]
    kind = rng.random(len(ctr)) < 0.5
    rad = rng.uniform(0.2, 1.5, len(ctr))
    hgt = rng.uniform(0.5, 4.0, len(ctr))
    for i in range(0, len(d), 16384):
        dd = d[i:i + 16384]
        best = t[i:i + 16384].copy()
        for j in range(len(ctr)):
            cx, cy, r, h = ctr[j, 0], ctr[j, 1], rad[j], hgt[j]
            if kind[j]:      # cylinder x^2+y^2 = r^2 around (cx, cy)
                a = dd[:, 0] ** 2 + dd[:, 1] ** 2
                b = -2 * (dd[:, 0] * cx + dd[:, 1] * cy)
                c = cx * cx + cy * cy - r * r
                disc = b * b - 4 * a * c
                ok = (disc >= 0) & (a > 1e-12)
                tt = np.where(ok, (-b - np.sqrt(np.maximum(disc, 0))) / (2 * np.maximum(a, 1e-12)), np.inf)
            else:            # box [cx-r, cx+r] x [cy-r, cy+r] x [-1.7, -1.7+h]
                lo = np.array([cx - r, cy - r, -1.7])
                hi = np.array([cx + r, cy + r, -1.7 + h])
                with np.errstate(divide="ignore", invalid="ignore"):
                    t1 = (lo - sensor_z * np.array([0, 0, 1])) / dd
                    t2 = (hi - sensor_z * np.array([0, 0, 1])) / dd
                tmin = np.nanmax(np.minimum(t1, t2), axis=1)
                tmax = np.nanmin(np.maximum(t1, t2), axis=1)
                tt = np.where((tmax >= tmin) & (tmax > 0), np.maximum(tmin, 0), np.inf)
            z = sensor_z + tt * dd[:, 2]
            valid = (tt > 0) & (z >= -1.7) & (z <= -1.7 + h)
            best = np.where(valid & (tt < best), tt, best)
        t[i:i + 16384] = best
    hit = np.isfinite(t) & (t <= rmax) & (t > 0.5)
    rr = t[hit] + rng.normal(0, 0.02, hit.sum())
    pts = d[hit] * rr[:, None] + np.array([0, 0, sensor_z])
    keep = voxel_first(pts, voxel)
    return pts[keep].astype(np.float32)


def subsample_to(xyz: np.ndarray, n: int, rng) -> np.ndarray:
    if len(xyz) <= n:
        return xyz
    sel = np.sort(rng.choice(len(xyz), n, replace=False))
    return xyz[sel]


# LiDAR sweep sizes after 5 cm voxel subsampling, by azimuth steps (64
# beams; measured on several seeds, +-5%): a cloud of n points is generated
# at the smallest step count expected to give about 1.1 n, then subsampled.
_LIDAR_STEPS = ((64, 3700), (128, 7500), (256, 15000), (512, 29000),
                (1024, 54000), (1536, 73000), (2048, 88000), (3072, 112000),
                (4096, 129000), (6144, 150000))


def lidar_points(n, seed):
    """A LiDAR sweep (``lidar``) of seed ``seed`` with exactly ``n`` points:
    the azimuth resolution is chosen to give at least n points after voxel
    subsampling (raised until it does), then n of them are kept at random."""
    i = next((j for j, (_, c) in enumerate(_LIDAR_STEPS) if c >= 1.1 * n),
             len(_LIDAR_STEPS) - 1)
    while True:
        pts = lidar(seed=seed, steps=_LIDAR_STEPS[i][0])
        if len(pts) >= n or i == len(_LIDAR_STEPS) - 1:
            break
        i += 1
    return subsample_to(pts, n, np.random.default_rng(seed + 10_000))


def room_points(n, seed):
    """A room (``room``) of seed ``seed`` with exactly ``n`` points (calibrated
    to about 1.05 n after voxel subsampling, then n kept at random)."""
    pts = room(seed=seed, target=int(n * 1.05) + 16)
    return subsample_to(pts, n, np.random.default_rng(seed + 10_000))


def batch(seed=50, n_clouds=16, lo=2000, hi=120_000):
    """(list of clouds, list of k): the uneven batched workload.  Sizes are
    exp(U(ln lo, ln hi)) and k is drawn from {8, 16, 32} (numpy generator
    ``seed``); cloud i comes from its own generator seeded ``seed + i``, the
    room generator for even i and the LiDAR generator for odd i."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_clouds)).astype(int)
    ks = rng.choice([8, 16, 32], n_clouds)
    clouds = []
    for i, n in enumerate(sizes):
        make = room_points if i % 2 == 0 else lidar_points
        clouds.append(make(int(n), seed + i))
    return clouds, [int(k) for k in ks]


def pack(clouds):
    """Concatenate clouds -> (xyz [N,3], int32 cumulative end offsets)."""
    xyz = np.concatenate(clouds, 0).astype(np.float32)
    off = np.cumsum([len(c) for c in clouds]).astype(np.int32)
    return xyz, off


def hierarchy(seed=51, sizes=(0.125, 0.25, 0.5)):
    """Room of ~100k points and its voxel-first subsampled stages
    (~100k -> ~25k -> ~6k -> ~1.5k)."""
    base = room(seed=seed)
    stages = [base]
    for s in sizes:
        stages.append(stages[-1][voxel_first(stages[-1], s)])
    return stages
