"""Seeded synthetic microstructures (fluid masks, 1 = fluid) for the permeability workloads.

All generators are periodic on an n^3 torus and deterministic for a given seed
(numpy PCG64).  Masks are uint8 arrays of shape (nz, ny, nx) == flat index
(k*ny + j)*nx + i, the reference's layout.
"""
from __future__ import annotations

import math

import numpy as np


def sphere_cell(n: int, radius_fraction: float = 0.3) -> np.ndarray:
    """Simple cubic array cell: voxel (i,j,k) is solid iff |(i+1/2, j+1/2, k+1/2) - n/2| <= r, r = f*n."""
    c = np.arange(n) + 0.5 - n / 2.0
    r = radius_fraction * n
    d2 = c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2
    return (d2 > r * r).astype(np.uint8)


def _ball_offsets(radius: float) -> np.ndarray:
    R = int(math.floor(radius))
    o = np.arange(-R, R + 1)
    dz, dy, dx = np.meshgrid(o, o, o, indexing="ij")
    keep = dx * dx + dy * dy + dz * dz <= radius * radius
    return np.stack([dz[keep], dy[keep], dx[keep]], axis=1)


def _stamp(solid: np.ndarray, centres: np.ndarray, offsets: np.ndarray, chunk: int = 256) -> None:
    n = solid.shape[0]
    flat = solid.reshape(-1)
    for s in range(0, len(centres), chunk):
        c = centres[s:s + chunk]
        pts = (c[:, None, :] + offsets[None, :, :]) % n
        idx = (pts[..., 0] * n + pts[..., 1]) * n + pts[..., 2]
        flat[idx.reshape(-1)] = 1


def boolean_spheres(n: int, radius: float, porosity: float, seed: int) -> np.ndarray:
    """Boolean model of overlapping lattice balls: Poisson(lambda N) integer centres, lambda = -ln(phi)/|B_r|."""
    rng = np.random.Generator(np.random.PCG64(seed))
    off = _ball_offsets(radius)
    lam = -math.log(porosity) / len(off)
    count = rng.poisson(lam * n ** 3)
    centres = rng.integers(0, n, size=(count, 3))
    solid = np.zeros((n, n, n), dtype=np.uint8)
    _stamp(solid, centres, off)
    return (1 - solid).astype(np.uint8)


def _ellipsoid_offsets(semi_axes, rot: np.ndarray) -> np.ndarray:
    a = max(semi_axes)
    R = int(math.ceil(a))
    o = np.arange(-R, R + 1)
    dz, dy, dx = np.meshgrid(o, o, o, indexing="ij")
    p = np.stack([dx.ravel(), dy.ravel(), dz.ravel()], axis=1).astype(np.float64)
    q = p @ rot  # body coordinates
    s = np.asarray(semi_axes, dtype=np.float64)
    keep = ((q / s) ** 2).sum(axis=1) <= 1.0
    pk = p[keep].astype(np.int64)
    return np.stack([pk[:, 2], pk[:, 1], pk[:, 0]], axis=1)


def _random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def boolean_ellipsoids(n: int, semi_axes=(12.0, 4.0, 4.0), porosity: float = 0.5, seed: int = 0) -> np.ndarray:
    """Boolean model of prolate ellipsoids with uniformly random orientations.  Per ellipsoid the
    generator draws the integer centre, then the rotation; the voxel sets are evaluated in
    batches (the union does not depend on the order)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    vol = 4.0 / 3.0 * math.pi * semi_axes[0] * semi_axes[1] * semi_axes[2]
    lam = -math.log(porosity) / vol
    count = rng.poisson(lam * n ** 3)
    centres = np.empty((count, 3), dtype=np.int64)
    rots = np.empty((count, 3, 3))
    for e in range(count):
        centres[e] = rng.integers(0, n, size=(1, 3))[0]
        rots[e] = _random_rotation(rng)
    solid = np.zeros((n, n, n), dtype=np.uint8)
    flat = solid.reshape(-1)
    R = int(math.ceil(max(semi_axes)))
    o = np.arange(-R, R + 1)
    dz, dy, dx = np.meshgrid(o, o, o, indexing="ij")
    p = np.stack([dx.ravel(), dy.ravel(), dz.ravel()], axis=1).astype(np.float64)
    pzyx = np.stack([dz.ravel(), dy.ravel(), dx.ravel()], axis=1).astype(np.int64)
    sa = np.asarray(semi_axes, dtype=np.float64)
    for s0 in range(0, count, 32):
        rb = rots[s0:s0 + 32]
        q = np.matmul(p[None], rb)  # body coordinates, p @ rot per ellipsoid
        keep = ((q / sa) ** 2).sum(axis=2) <= 1.0
        pts = (centres[s0:s0 + 32, None, :] + pzyx[None])[keep] % n
        flat[(pts[:, 0] * n + pts[:, 1]) * n + pts[:, 2]] = 1
    return (1 - solid).astype(np.uint8)


def slit_channel(n: int, width: int) -> np.ndarray:
    """Plane channel: fluid for z in [0, width), solid elsewhere (periodic in x, y)."""
    m = np.zeros((n, n, n), dtype=np.uint8)
    m[:width] = 1
    return m
