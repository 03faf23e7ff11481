"""Seeded synthetic point clouds for the pipeline's workloads.

float4 points with w = 0 (the reference layout, include/shared/types.h), or
float3, coordinates clipped to [0, nextafter(1024, 0)] so the Morton
quantisation never wraps (include/shared/morton.h:42-44 maps 1024 to 0).

- ``uniform``: each coordinate uniform in [0, 1024) (the reference's own
  generator draws the same distribution, kernels/00_init.cpp:20-30).
- ``clustered``: a surface-like scene with heavy duplicate Morton keys:
  40% on a z = 0 ground plane with N(0, 0.5) jitter, 40% on the surfaces of
  32 vertical cylinders (radius 2-20, height 5-200), 20% in 64 isotropic
  Gaussians (sigma 1-30).
- ``batch``: clouds with log-uniform sizes, each uniform or clustered at
  random, each scaled and shifted so that it has its own bounding box, plus
  the per-cloud cube box (min, range) the pipeline quantises with.
- ``frames``: successive small rigid motions of one cloud.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

LIMIT = np.nextafter(np.float32(1024.0), np.float32(0.0))


def _pack(xyz: np.ndarray, float4: bool) -> np.ndarray:
    xyz = np.clip(xyz, 0.0, LIMIT).astype(np.float32)
    if not float4:
        return np.ascontiguousarray(xyz)
    out = np.zeros((xyz.shape[0], 4), np.float32)
    out[:, :3] = xyz
    return out


def uniform(n: int, seed: int = 114514, float4: bool = True) -> np.ndarray:
    rng = np.random.default_rng(seed)
    xyz = rng.random((n, 3), dtype=np.float32) * np.float32(1024.0)
    return _pack(xyz, float4)


def clustered(n: int, seed: int = 114514, float4: bool = True) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n_plane = int(round(0.4 * n))
    n_cyl = int(round(0.4 * n))
    n_gauss = n - n_plane - n_cyl
    xyz = np.empty((n, 3), np.float32)
    # ground plane
    xyz[:n_plane, 0:2] = rng.random((n_plane, 2), dtype=np.float32) * 1024.0
    xyz[:n_plane, 2] = rng.normal(0.0, 0.5, n_plane)
    # 32 vertical cylinder surfaces
    cen = rng.random((32, 2)) * 1024.0
    rad = rng.uniform(2.0, 20.0, 32)
    hgt = rng.uniform(5.0, 200.0, 32)
    which = rng.integers(0, 32, n_cyl)
    theta = rng.random(n_cyl) * (2.0 * np.pi)
    sl = slice(n_plane, n_plane + n_cyl)
    xyz[sl, 0] = cen[which, 0] + rad[which] * np.cos(theta)
    xyz[sl, 1] = cen[which, 1] + rad[which] * np.sin(theta)
    xyz[sl, 2] = rng.random(n_cyl) * hgt[which]
    # 64 isotropic Gaussians
    gc = rng.random((64, 3)) * 1024.0
    gs = rng.uniform(1.0, 30.0, 64)
    which = rng.integers(0, 64, n_gauss)
    xyz[n_plane + n_cyl:] = gc[which] + rng.normal(0.0, 1.0, (n_gauss, 3)) * gs[which, None]
    perm = rng.permutation(n)
    return _pack(xyz[perm], float4)


def cube_box(pts: np.ndarray):
    """(min, range) of a cube enclosing the cloud, with the far face mapped
    strictly below 1024 quanta."""
    xyz = pts[:, :3]
    mn = np.float32(xyz.min()) if xyz.size else np.float32(0.0)
    ext = np.float32(xyz.max() - mn) if xyz.size else np.float32(1.0)
    rg = np.float32(max(float(ext) * 1.0001, 1e-3))
    return float(mn), float(rg)


def batch(n_clouds: int = 512, lo: int = 4096, hi: int = 2 * 1024 * 1024,
          seed: int = 114514, float4: bool = True, max_total: int | None = None,
          workers: int | None = None):
    """Returns (points, seg_offsets[n+1], seg_min[n], seg_range[n], kinds).
    Clouds are generated on `workers` threads (each from its own seed; the
    batch's own draws come first), so the result does not depend on it."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_clouds)).astype(np.int64)
    if max_total is not None and sizes.sum() > max_total:
        sizes = np.maximum((sizes * (max_total / sizes.sum())).astype(np.int64), 1)
    kinds = rng.integers(0, 2, n_clouds)
    scales = np.zeros(n_clouds)
    shifts = np.zeros((n_clouds, 3))
    for k in range(n_clouds):
        scales[k] = rng.uniform(0.05, 4.0)
        shifts[k] = rng.uniform(-2000.0, 2000.0, 3)
    offs = np.zeros(n_clouds + 1, np.int64)
    offs[1:] = np.cumsum(sizes)
    width = 4 if float4 else 3
    pts = np.zeros((int(offs[-1]), width), np.float32)
    smin = np.zeros(n_clouds, np.float32)
    srg = np.zeros(n_clouds, np.float32)

    def one(k: int) -> None:
        gen = clustered if kinds[k] else uniform
        c = gen(int(sizes[k]), seed=seed * 7919 + k, float4=False).astype(np.float64)
        c = (c * scales[k] + shifts[k]).astype(np.float32)
        pts[offs[k]:offs[k + 1], :3] = c
        smin[k], srg[k] = cube_box(c)

    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    if workers <= 1:
        for k in range(n_clouds):
            one(k)
    else:
        with ThreadPoolExecutor(max_workers=workers) as ex:
            list(ex.map(one, range(n_clouds)))
    return pts, offs.astype(np.int32), smin, srg, kinds


def rigid_motion(pts: np.ndarray, seed: int, max_deg: float = 0.5,
                 max_shift: float = 2.0) -> np.ndarray:
    """Rotate about the cloud centre by <= max_deg about a random axis and
    shift by <= max_shift per axis; clipped back into the box."""
    rng = np.random.default_rng(seed)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    ang = np.deg2rad(rng.uniform(-max_deg, max_deg))
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    rot = np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * (k @ k)
    shift = rng.uniform(-max_shift, max_shift, 3)
    xyz = pts[:, :3].astype(np.float64)
    c = xyz.mean(axis=0)
    moved = (xyz - c) @ rot.T + c + shift
    out = pts.copy()
    out[:, :3] = np.clip(moved, 0.0, LIMIT).astype(np.float32)
    return out
