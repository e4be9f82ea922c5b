"""FDK / FBP normalisation, pinned on the CPU: analytic projections of a uniform
object, weighted and Ram-Lak filtered with numpy (np.convolve), backprojected by
the oracle with the plan's multiplier, must give back the density (textbook
FBP; no reference implementation exists for this row)."""
import math

import numpy as np
import pytest

import oracle as orc
from pyronn_layers import geometry as geo
from pyronn_layers import phantoms as ph


def ramlak(tau, n):
    k = np.arange(-(n - 1), n)
    h = np.zeros(k.shape)
    h[k == 0] = 1.0 / (4 * tau * tau)
    odd = (k % 2) == 1
    h[odd] = -1.0 / (k[odd] ** 2 * math.pi ** 2 * tau * tau)
    return h


def filter_rows(rows, tau):
    n = rows.shape[-1]
    h = ramlak(tau, n)
    out = np.empty_like(rows, dtype=np.float64)
    flat = rows.reshape(-1, n)
    o = out.reshape(-1, n)
    for i in range(flat.shape[0]):
        o[i] = tau * np.convolve(flat[i], h)[n - 1:2 * n - 1]
    return out


def _ball(radius_mm, density=1.0):
    return ph.Ellipsoids(centres=np.zeros((1, 3)), axes=np.full((1, 3), radius_mm), rot=np.eye(3)[None],
                         density=np.array([density]))


def test_cone_fdk_uniform_ball():
    D = H = W = 32
    sp = np.ones(3)
    og = geo.centred_origin((D, H, W), sp)
    n, Hd, Wd, du = 120, 48, 64, 1.0
    sid, sdd = 120.0, 200.0
    P = geo.circular_projection_matrices(n, sid, sdd, (Hd, Wd), (du, du))
    e = _ball(10.0, 0.8)
    vv, uu = np.meshgrid(np.arange(Hd), np.arange(Wd), indexing="ij")
    sino = np.empty((n, Hd, Wd))
    for k in range(n):
        M, p4 = P[k][:, :3], P[k][:, 3]
        S = -np.linalg.solve(M, p4)
        dirs = np.linalg.solve(M, np.stack([uu, vv, np.ones_like(uu)], 0).reshape(3, -1)).T
        sino[k] = ph.ellipsoid_line_integrals(e, np.broadcast_to(S, dirs.shape), dirs).reshape(Hd, Wd)
    a = (uu - (Wd - 1) / 2) * du
    b = (vv - (Hd - 1) / 2) * du
    cw = sdd / np.sqrt(sdd ** 2 + a ** 2 + b ** 2)
    q = filter_rows(sino * cw[None], du * sid / sdd)
    mult = math.pi * sid ** 2 / n
    rec, _ = orc.cone_backward(q.astype(np.float32), P.astype(np.float32), (D, H, W), sp.astype(np.float32),
                               og.astype(np.float32), mult=mult)
    c = D // 2
    core = rec[c - 3:c + 3, c - 5:c + 5, c - 5:c + 5]
    assert core.mean() == pytest.approx(0.8, rel=0.03)
    outside = rec[c, :3, :3]
    assert abs(outside.mean()) < 0.05


def test_fan_fbp_uniform_disc():
    H = W = 64
    s = (0.5, 0.5)
    og = geo.centred_origin((H, W), s)
    n, det, ds = 180, 129, 0.5
    do = -(det - 1) / 2 * ds
    sid, sdd = 60.0, 100.0
    rays = geo.fan_central_ray_vectors(n, 2 * math.pi)
    e = ph.Ellipses(centres=np.zeros((1, 2)), axes=np.full((1, 2), 8.0), angle=np.zeros(1), density=np.array([1.0]))
    u = do + ds * np.arange(det)
    r = rays[:, None, :]
    uv = np.stack([-rays[:, 1], rays[:, 0]], 1)[:, None, :]
    src = -sid * r
    detp = src + sdd * r + uv * u[None, :, None]
    sino = ph.ellipse_line_integrals(e, np.broadcast_to(src, detp.shape), detp - src)
    cw = sdd / np.sqrt(sdd ** 2 + u ** 2)
    q = filter_rows(sino * cw[None], ds * sid / sdd)
    q *= ((sdd ** 2 + u ** 2) / sdd ** 2 * sid / sdd)[None]
    rec, _ = orc.fan_backward(q.astype(np.float32), rays.astype(np.float32), (H, W), og, s, do, ds, sid, sdd)
    c = H // 2
    assert rec[c - 6:c + 6, c - 6:c + 6].mean() == pytest.approx(1.0, rel=0.02)
    assert abs(rec[:4, :4].mean()) < 0.05


@pytest.mark.parametrize("rng_", [math.pi, 2 * math.pi])
def test_parallel_fbp_uniform_disc(rng_):
    H = W = 64
    s = (1.0, 1.0)
    og = geo.centred_origin((H, W), s)
    n, det, ds = 120, 95, 1.0
    do = -(det - 1) / 2 * ds
    rays = geo.parallel_ray_vectors(n, rng_)
    e = ph.Ellipses(centres=np.zeros((1, 2)), axes=np.full((1, 2), 18.0), angle=np.zeros(1), density=np.array([1.0]))
    u = do + ds * np.arange(det)
    uv = np.stack([-rays[:, 1], rays[:, 0]], 1)[:, None, :]
    sino = ph.ellipse_line_integrals(e, uv * u[None, :, None], np.broadcast_to(rays[:, None, :], (n, det, 2)))
    q = filter_rows(sino, ds)
    rec, _ = orc.parallel_backward(q.astype(np.float32), rays.astype(np.float32), (H, W), og, s, do, ds)
    c = H // 2
    assert rec[c - 8:c + 8, c - 8:c + 8].mean() == pytest.approx(1.0, rel=0.02)
