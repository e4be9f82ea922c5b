"""Work counts for the throughput metrics.

Forward "samples" are the distinct alpha positions of the reference march
(cone_projector_3D_CudaKernel.cu.cc:135-178): the entrance sample, every loop
sample and the exit sample, for each ray that enters the clip box.  The count
depends only on geometry.  Here it is evaluated in closed form in float64,
K_loop = ceil((alpha_max - alpha_0) / h), which equals the float32 march's
count except for rays whose alpha lands within rounding of alpha_max (a
relative difference of order 1e-7 on a full problem; checked against the
oracle's exact count in the tests).
Backprojection "updates" are (voxel, view) pairs: W*H*D*n.
"""
from __future__ import annotations

import numpy as np


def cone_forward_sample_count(P, volume_shape, spacing_zyx, origin_zyx, detector_shape, step,
                              hardware_interp=False, device=None) -> int:
    """Exact-up-to-ties number of forward samples for one cone projection."""
    import torch

    dev = torch.device(device) if device is not None else torch.device("cuda" if torch.cuda.is_available() else "cpu")
    f64 = torch.float64
    P = np.asarray(P, np.float64).reshape(-1, 3, 4)
    D, H, W = volume_shape
    Hd, Wd = detector_shape
    sp = np.asarray(spacing_zyx, np.float64)[::-1].copy()
    og = np.asarray(origin_zyx, np.float64)[::-1].copy()
    lo = -0.5 if hardware_interp else 0.0
    hi = np.array([W, H, D], np.float64) - (0.5 if hardware_interp else 0.0)
    vv, uu = torch.meshgrid(torch.arange(Hd, device=dev, dtype=f64), torch.arange(Wd, device=dev, dtype=f64),
                            indexing="ij")
    pix = torch.stack([uu.reshape(-1), vv.reshape(-1), torch.ones_like(uu).reshape(-1)], 0)
    total = 0
    for k in range(P.shape[0]):
        M, p4 = P[k][:, :3], P[k][:, 3]
        Minv = np.linalg.inv(M)
        C = -Minv @ p4
        s = torch.tensor((C - og) / sp, device=dev, dtype=f64)
        A = torch.tensor(Minv / sp[:, None], device=dev, dtype=f64)
        r = A @ pix
        r = r / torch.linalg.norm(r, dim=0, keepdim=True)
        amin = torch.full_like(r[0], -float("inf"))
        amax = torch.full_like(r[0], float("inf"))
        for a in range(3):
            ra = r[a]
            nz = ra != 0
            inv = torch.where(nz, 1.0 / torch.where(nz, ra, torch.ones_like(ra)), torch.zeros_like(ra))
            a0 = (lo - s[a]) * inv
            a1 = (hi[a] - s[a]) * inv
            mn = torch.where(nz, torch.minimum(a0, a1), torch.full_like(ra, -float("inf")))
            mx = torch.where(nz, torch.maximum(a0, a1), torch.full_like(ra, float("inf")))
            if a == 0:
                amin = torch.where(nz, mn, torch.zeros_like(ra))
            else:
                amin = torch.maximum(amin, mn)
            amax = torch.minimum(amax, mx)
        a0 = amin + 0.5 * step
        hit = a0 < amax
        loop = torch.clamp(torch.ceil((amax - (a0 + step)) / step), min=0)
        n = torch.where(hit, loop + 2, torch.zeros_like(loop))
        total += int(n.sum().item())
    return total


def cone_backprojection_update_count(volume_shape, n_views) -> int:
    D, H, W = volume_shape
    return int(D) * int(H) * int(W) * int(n_views)
