"""Seeded ellipse / ellipsoid phantoms and their analytic line integrals.

Phantoms are sums of uniform-density ellipses (2D) or ellipsoids (3D) placed in
world millimetres.  Parameters come from numpy.random.default_rng(seed), so a
seed fixes the phantom exactly.  Rasterisation samples each voxel centre; the
analytic line integrals (chord length x density) give a discretisation-level
check of the projectors.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

import numpy as np


@dataclass
class Ellipsoids:
    centres: np.ndarray   # [k, 3] (x, y, z) mm
    axes: np.ndarray      # [k, 3] semi-axes mm
    rot: np.ndarray       # [k, 3, 3] columns = principal directions (body coords = p @ rot)
    density: np.ndarray   # [k]


@dataclass
class Ellipses:
    centres: np.ndarray   # [k, 2] (x, y) mm
    axes: np.ndarray      # [k, 2]
    angle: np.ndarray     # [k]
    density: np.ndarray   # [k]


def _haar_rotation(rng: np.random.Generator) -> np.ndarray:
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def random_ellipsoids(seed: int, count: int, volume_shape: Sequence[int], spacing: Sequence[float],
                      axes_range_vox=(8.0, 100.0)) -> Ellipsoids:
    """`count` ellipsoids with semi-axes U[axes_range] voxels, Haar-random rotation,
    density U[0,1], each kept inside the sphere inscribed in the volume."""
    rng = np.random.default_rng(seed)
    D, H, W = volume_shape
    sz, sy, sx = spacing
    vox = min(sx, sy, sz)
    radius = 0.5 * min(W * sx, H * sy, D * sz)
    lo, hi = axes_range_vox
    hi = min(hi, 0.95 * radius / vox)
    lo = min(lo, hi)
    centres, axes, rots, dens = [], [], [], []
    for _ in range(count):
        a = rng.uniform(lo, hi, size=3) * vox
        R = _haar_rotation(rng)
        d = rng.uniform(0.0, 1.0)
        amax = a.max()
        room = max(radius - amax, 0.0)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u) + 1e-300
        c = u * room * rng.uniform(0.0, 1.0) ** (1.0 / 3.0)
        centres.append(c)
        axes.append(a)
        rots.append(R)
        dens.append(d)
    return Ellipsoids(np.array(centres), np.array(axes), np.array(rots), np.array(dens))


def random_ellipses(seed: int, count: int, volume_shape: Sequence[int], spacing: Sequence[float],
                    axes_range_px=(10.0, 90.0)) -> Ellipses:
    rng = np.random.default_rng(seed)
    H, W = volume_shape
    sy, sx = spacing
    px = min(sx, sy)
    radius = 0.5 * min(W * sx, H * sy)
    lo, hi = axes_range_px
    hi = min(hi, 0.95 * radius / px)
    lo = min(lo, hi)
    centres, axes, ang, dens = [], [], [], []
    for _ in range(count):
        a = rng.uniform(lo, hi, size=2) * px
        t = rng.uniform(0.0, np.pi)
        d = rng.uniform(0.0, 1.0)
        room = max(radius - a.max(), 0.0)
        phi = rng.uniform(0, 2 * np.pi)
        rr = room * np.sqrt(rng.uniform())
        centres.append([rr * np.cos(phi), rr * np.sin(phi)])
        axes.append(a)
        ang.append(t)
        dens.append(d)
    return Ellipses(np.array(centres), np.array(axes), np.array(ang), np.array(dens))


def rasterise_ellipsoids(e: Ellipsoids, volume_shape, spacing, origin, dtype=np.float32) -> np.ndarray:
    """Voxel-centre rasterisation into [D, H, W] (origin/spacing in z,y,x order)."""
    D, H, W = volume_shape
    sz, sy, sx = spacing
    oz, oy, ox = origin
    x = ox + sx * np.arange(W)
    y = oy + sy * np.arange(H)
    vol = np.zeros((D, H, W), np.float64)
    for k in range(len(e.density)):
        R, a, c = e.rot[k], e.axes[k], e.centres[k]
        zs = oz + sz * np.arange(D)
        # Restrict to the bounding slab in z for speed.
        ext = np.abs(R) @ a
        zsel = np.nonzero(np.abs(zs - c[2]) <= ext[2] + sz)[0]
        if zsel.size == 0:
            continue
        Z, Y, X = np.meshgrid(zs[zsel], y, x, indexing="ij")
        P = np.stack([X - c[0], Y - c[1], Z - c[2]], axis=-1)
        B = P @ R  # body coordinates (columns of R are principal axes)
        q = (B[..., 0] / a[0]) ** 2 + (B[..., 1] / a[1]) ** 2 + (B[..., 2] / a[2]) ** 2
        vol[zsel] += e.density[k] * (q <= 1.0)
    return vol.astype(dtype)


def rasterise_ellipses(e: Ellipses, volume_shape, spacing, origin, dtype=np.float32) -> np.ndarray:
    H, W = volume_shape
    sy, sx = spacing
    oy, ox = origin
    X, Y = np.meshgrid(ox + sx * np.arange(W), oy + sy * np.arange(H), indexing="xy")
    img = np.zeros((H, W), np.float64)
    for k in range(len(e.density)):
        ct, st = np.cos(e.angle[k]), np.sin(e.angle[k])
        dx, dy = X - e.centres[k, 0], Y - e.centres[k, 1]
        bx = ct * dx + st * dy
        by = -st * dx + ct * dy
        img += e.density[k] * ((bx / e.axes[k, 0]) ** 2 + (by / e.axes[k, 1]) ** 2 <= 1.0)
    return img.astype(dtype)


def ellipsoid_line_integrals(e: Ellipsoids, src: np.ndarray, dirs: np.ndarray) -> np.ndarray:
    """Sum over ellipsoids of density x chord length for lines src + t*dir (world mm).

    src [..., 3], dirs [..., 3] (any length; normalised here)."""
    src = np.asarray(src, np.float64)
    d = np.asarray(dirs, np.float64)
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    out = np.zeros(np.broadcast_shapes(src.shape, d.shape)[:-1])
    for k in range(len(e.density)):
        R, a, c = e.rot[k], e.axes[k], e.centres[k]
        p = (src - c) @ R / a
        q = d @ R / a
        A = np.sum(q * q, -1)
        B = 2 * np.sum(p * q, -1)
        C = np.sum(p * p, -1) - 1.0
        disc = B * B - 4 * A * C
        chord = np.where(disc > 0, np.sqrt(np.maximum(disc, 0)) / A, 0.0)
        out += e.density[k] * chord
    return out


def ellipse_line_integrals(e: Ellipses, src: np.ndarray, dirs: np.ndarray) -> np.ndarray:
    src = np.asarray(src, np.float64)
    d = np.asarray(dirs, np.float64)
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    out = np.zeros(np.broadcast_shapes(src.shape, d.shape)[:-1])
    for k in range(len(e.density)):
        ct, st = np.cos(e.angle[k]), np.sin(e.angle[k])
        Rm = np.array([[ct, st], [-st, ct]])
        p = ((src - e.centres[k]) @ Rm.T) / e.axes[k]
        q = (d @ Rm.T) / e.axes[k]
        A = np.sum(q * q, -1)
        B = 2 * np.sum(p * q, -1)
        C = np.sum(p * p, -1) - 1.0
        disc = B * B - 4 * A * C
        out += e.density[k] * np.where(disc > 0, np.sqrt(np.maximum(disc, 0)) / A, 0.0)
    return out


def rasterise_ellipsoids_torch(e: Ellipsoids, volume_shape, spacing, origin, device="cuda",
                               dtype=None, slab: int = 128):
    """Same rasterisation as rasterise_ellipsoids, computed with torch on `device`
    (for full-size phantoms).  Accumulates in float64, returns float32."""
    import torch

    D, H, W = volume_shape
    sz, sy, sx = spacing
    oz, oy, ox = origin
    f64 = torch.float64
    x = ox + sx * torch.arange(W, device=device, dtype=f64)
    y = oy + sy * torch.arange(H, device=device, dtype=f64)
    vol = torch.zeros((D, H, W), device=device, dtype=torch.float32)
    for z0 in range(0, D, slab):
        z1 = min(D, z0 + slab)
        zs = oz + sz * torch.arange(z0, z1, device=device, dtype=f64)
        Z, Y, X = torch.meshgrid(zs, y, x, indexing="ij")
        acc = torch.zeros_like(X)
        for k in range(len(e.density)):
            R = torch.tensor(e.rot[k], device=device, dtype=f64)
            a = e.axes[k]
            c = e.centres[k]
            dx, dy, dz = X - c[0], Y - c[1], Z - c[2]
            b0 = dx * R[0, 0] + dy * R[1, 0] + dz * R[2, 0]
            b1 = dx * R[0, 1] + dy * R[1, 1] + dz * R[2, 1]
            b2 = dx * R[0, 2] + dy * R[1, 2] + dz * R[2, 2]
            q = (b0 / a[0]) ** 2 + (b1 / a[1]) ** 2 + (b2 / a[2]) ** 2
            acc += float(e.density[k]) * (q <= 1.0).to(f64)
        vol[z0:z1] = acc.to(torch.float32)
    return vol if dtype is None else vol.to(dtype)
