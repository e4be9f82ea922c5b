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
    Minv = np.linalg.inv(P[:, :, :3])
    C = -np.einsum("nij,nj->ni", Minv, P[:, :, 3])
    S_all = torch.tensor((C - og) / sp, device=dev, dtype=f64)            # [n, 3]
    A_all = torch.tensor(Minv / sp[None, :, None], device=dev, dtype=f64)  # [n, 3, 3]
    chunk = max(1, int(2 ** 26 // max(1, Hd * Wd)))
    total = 0
    inf = float("inf")
    for k0 in range(0, P.shape[0], chunk):
        A = A_all[k0:k0 + chunk]
        s = S_all[k0:k0 + chunk, :, None]
        r = A @ pix                                   # [c, 3, R]
        r = r / torch.linalg.norm(r, dim=1, keepdim=True)
        nz = r != 0
        inv = 1.0 / torch.where(nz, r, torch.ones_like(r))
        hi_t = torch.tensor(hi, device=dev, dtype=f64)[None, :, None]
        a0 = (lo - s) * inv
        a1 = (hi_t - s) * inv
        mn = torch.where(nz, torch.minimum(a0, a1), torch.full_like(r, -inf))
        mx = torch.where(nz, torch.maximum(a0, a1), torch.full_like(r, inf))
        # the x slab overwrites the initial [0, inf); y and z intersect with it
        amin = torch.where(nz[:, 0], mn[:, 0], torch.zeros_like(mn[:, 0]))
        amin = torch.maximum(amin, torch.maximum(mn[:, 1], mn[:, 2]))
        amax = torch.minimum(mx[:, 0], torch.minimum(mx[:, 1], mx[:, 2]))
        first = amin + 0.5 * step
        hit = first < amax
        loop = torch.clamp(torch.ceil((amax - (first + step)) / step), min=0)
        total += int(torch.where(hit, loop + 2, torch.zeros_like(loop)).sum().item())
        del r, nz, inv, a0, a1, mn, mx
    return total


def cone_backprojection_update_count(volume_shape, n_views) -> int:
    D, H, W = volume_shape
    return int(D) * int(H) * int(W) * int(n_views)
