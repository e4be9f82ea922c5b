"""Trajectory builders producing the geometry inputs of the six operators.

The operators take their geometry as arrays (cc/ops/*_OPKernel.cc REGISTER_OP
input lists): 2D ops take per-view ray vectors (x, y); cone ops take per-view
3x4 projection matrices P that map world millimetres (x, y, z, 1) to
(u*w, v*w, w), u = detector column, v = detector row, pixel centres at integer
(u, v) (cone_projector_3D_CudaKernel.cu.cc:200-202,
cone_backprojector_3D_CudaKernel.cu.cc:105-110).  Volume origin and spacing
arrays are stored (z, y, x) / (y, x) (cone_projector_3D_CudaKernel.cu.cc:191).

All builders compute in float64; callers cast to float32 for the operators.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Sequence, Tuple

import numpy as np


def centred_origin(shape: Sequence[int], spacing: Sequence[float]) -> np.ndarray:
    """Origin that puts the volume centre at the world origin: -(N-1)/2 * s per axis."""
    shape = np.asarray(shape, np.float64)
    spacing = np.asarray(spacing, np.float64)
    return -(shape - 1.0) / 2.0 * spacing


def parallel_ray_vectors(n_views: int, angular_range: float = math.pi, start: float = 0.0) -> np.ndarray:
    """Unit ray directions (cos t, sin t) for n views evenly spaced over `angular_range`."""
    t = start + np.arange(n_views, dtype=np.float64) * (angular_range / n_views)
    return np.stack([np.cos(t), np.sin(t)], axis=1)


def fan_central_ray_vectors(n_views: int, angular_range: float = 2 * math.pi, start: float = 0.0) -> np.ndarray:
    """Central ray directions for a fan orbit; the source sits at -SID * ray."""
    return parallel_ray_vectors(n_views, angular_range, start)


def circular_projection_matrices(n_views: int, sid: float, sdd: float, detector_shape: Tuple[int, int],
                                 detector_spacing: Tuple[float, float], angular_range: float = 2 * math.pi,
                                 start: float = 0.0, principal_point: Tuple[float, float] | None = None
                                 ) -> np.ndarray:
    """P matrices [n, 3, 4] for a circular cone-beam orbit about the z axis.

    View k at angle b = start + k * range / n: source S = SID * (cos b, sin b, 0);
    central ray d = -(cos b, sin b, 0); detector column axis e_u = (-sin b, cos b, 0),
    row axis e_v = (0, 0, 1).  With R = [e_u; e_v; d] and
    K = [[SDD/du, 0, u0], [0, SDD/dv, v0], [0, 0, 1]], P = K R [I | -S], so w is the
    depth of the point along the central ray in mm and the isocentre maps to the
    principal point (u0, v0) = ((W_d-1)/2, (H_d-1)/2) by default.
    detector_shape is (H_d, W_d), detector_spacing (dv, du).
    """
    Hd, Wd = detector_shape
    dv, du = detector_spacing
    if principal_point is None:
        u0, v0 = (Wd - 1) / 2.0, (Hd - 1) / 2.0
    else:
        u0, v0 = principal_point
    K = np.array([[sdd / du, 0.0, u0], [0.0, sdd / dv, v0], [0.0, 0.0, 1.0]])
    out = np.empty((n_views, 3, 4), np.float64)
    for k in range(n_views):
        b = start + k * angular_range / n_views
        c, s = math.cos(b), math.sin(b)
        S = np.array([sid * c, sid * s, 0.0])
        R = np.array([[-s, c, 0.0], [0.0, 0.0, 1.0], [-c, -s, 0.0]])
        KR = K @ R
        out[k, :, :3] = KR
        out[k, :, 3] = -KR @ S
    return out


@dataclass
class ConeGeometry:
    """Circular cone-beam geometry in the reference's array conventions."""
    volume_shape: Tuple[int, int, int]          # (D, H, W)
    volume_spacing: Tuple[float, float, float]  # (z, y, x) mm
    detector_shape: Tuple[int, int]             # (H_d, W_d)
    detector_spacing: Tuple[float, float]       # (dv, du) mm
    n_views: int
    sid: float
    sdd: float
    angular_range: float = 2 * math.pi
    step_size: float = 0.5
    projection_multiplier: float = 1.0
    volume_origin: np.ndarray = field(default=None)  # (z, y, x) mm

    def __post_init__(self):
        if self.volume_origin is None:
            self.volume_origin = centred_origin(self.volume_shape, self.volume_spacing)
        self.volume_origin = np.asarray(self.volume_origin, np.float64)

    @property
    def projection_shape(self) -> Tuple[int, int, int]:
        return (self.n_views,) + tuple(self.detector_shape)

    def projection_matrices(self) -> np.ndarray:
        return circular_projection_matrices(self.n_views, self.sid, self.sdd, self.detector_shape,
                                            self.detector_spacing, self.angular_range)


@dataclass
class Geometry2D:
    """Parallel or fan 2D geometry in the reference's array conventions."""
    volume_shape: Tuple[int, int]            # (H, W)
    volume_spacing: Tuple[float, float]      # (y, x) mm
    detector_shape: int
    detector_spacing: float
    n_views: int
    angular_range: float
    sid: float = 0.0
    sdd: float = 0.0
    step_size: float = 0.2                   # x min(spacing); the reference hard-codes 0.2
    volume_origin: np.ndarray = field(default=None)
    detector_origin: float = None

    def __post_init__(self):
        if self.volume_origin is None:
            self.volume_origin = centred_origin(self.volume_shape, self.volume_spacing)
        self.volume_origin = np.asarray(self.volume_origin, np.float64)
        if self.detector_origin is None:
            self.detector_origin = -(self.detector_shape - 1) / 2.0 * self.detector_spacing

    @property
    def projection_shape(self) -> Tuple[int, int]:
        return (self.n_views, self.detector_shape)

    def ray_vectors(self) -> np.ndarray:
        return parallel_ray_vectors(self.n_views, self.angular_range)
