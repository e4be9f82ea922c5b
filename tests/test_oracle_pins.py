"""Pins for the CPU oracle: closed forms, library routines and invariants.

Each test is chosen so that a plausible restatement mistake (swapped axis,
wrong clip box, dropped tail term, wrong weight convention, transposed
matrix) fails it.  None of them re-types the oracle's own loop.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle as orc
from pyronn_layers import geometry as geo
from pyronn_layers import phantoms as ph


def _axis_ray(axis, sign=1.0):
    """inv_ar for a 1x1 detector whose only ray points along +/-axis (index space)."""
    inv = np.zeros((1, 9), np.float32)
    inv[0, 2 + 3 * axis] = sign
    return inv


# ----------------------------------------------------------------------------- setup (A1)

def test_cone_setup_matches_numpy_inverse_and_svd_nullspace():
    P = geo.circular_projection_matrices(7, 600.0, 1100.0, (48, 64), (1.3, 1.1), start=0.3).astype(np.float32)
    spacing = np.array([0.7, 1.2, 0.9], np.float32)   # z, y, x
    origin = np.array([-11.0, 5.0, -30.0], np.float32)
    inv, src, inv64, src64, st = orc.cone_setup(P, spacing, origin)
    assert st == 0
    for k in range(7):
        Pd = P[k].astype(np.float64)
        Minv = np.linalg.inv(Pd[:, :3])
        sxyz = spacing[::-1].astype(np.float64)
        np.testing.assert_allclose(inv64[k].reshape(3, 3), Minv / sxyz[:, None], rtol=1e-11, atol=1e-17)
        _, _, vt = np.linalg.svd(Pd)
        C = vt[-1] / vt[-1][3]
        np.testing.assert_allclose(src64[k], (C[:3] - origin[::-1]) / sxyz, rtol=1e-9, atol=1e-9)
    np.testing.assert_array_equal(inv, inv64.astype(np.float32))
    np.testing.assert_array_equal(src, src64.astype(np.float32))


def test_cone_setup_rejects_singular_matrix():
    P = np.zeros((2, 3, 4), np.float32)
    P[0] = geo.circular_projection_matrices(1, 500.0, 900.0, (8, 8), (1.0, 1.0))[0]
    P[1, :, :3] = [[1, 2, 3], [2, 4, 6], [0, 0, 1]]   # rank 2
    _, _, inv64, _, st = orc.cone_setup(P, np.ones(3, np.float32), np.zeros(3, np.float32))
    assert st == 1
    assert np.all(np.isfinite(inv64[0])) and np.all(np.isnan(inv64[1]))


# ----------------------------------------------------------------------------- cone forward (A3-A6)

SHAPE = (11, 9, 7)                                   # D, H, W
SPACING = np.array([2.0, 0.8, 1.5], np.float32)      # z, y, x


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_cone_forward_linear_ramp_is_exact(axis, sign):
    """A volume that falls linearly to zero at the far face of the [0,N] box is
    linear along an axis ray over the whole box, so the trapezoid + tail is exact:
    +dir: integral_{h/2}^{N} (N - x) dx;  -dir: integral_0^{N-h/2} (N - x) dx."""
    D, H, W = SHAPE
    N = (W, H, D)[axis]
    idx = np.indices(SHAPE)[[2, 1, 0][axis]].astype(np.float32)
    vol = (N - idx).astype(np.float32)
    src = np.array([[3.0, 4.0, 5.0]], np.float32)
    src[0, axis] = -10.0 if sign > 0 else N + 10.0
    h = 0.5
    out, _, ns = orc.cone_forward(vol, _axis_ray(axis, sign), src, (1, 1), SPACING, h)
    scale = SPACING[::-1][axis]
    if sign > 0:
        exact = (N - h / 2) ** 2 / 2
    else:
        b = N - h / 2
        exact = N * b - b * b / 2
    assert out[0, 0, 0] == pytest.approx(scale * exact, rel=2e-6)
    assert ns == 2 * N + 1   # entrance + loop samples + exit


def test_cone_forward_negative_volume_skips_tail():
    """Q6: the tail runs only for a positive sum.  alpha values are exact binary
    fractions here, so the plain trapezoid sum is known exactly."""
    D, H, W = SHAPE
    vol = -(W - np.indices(SHAPE)[2]).astype(np.float32)
    src = np.array([[-10.0, 4.0, 5.0]], np.float32)
    h = 0.5
    out, _, ns = orc.cone_forward(vol, _axis_ray(0), src, (1, 1), SPACING, h)
    a = h / 2 + h * np.arange(2 * W)
    f = -(W - a)
    expect = h * (0.5 * f[0] + f[1:].sum()) * SPACING[2]
    assert out[0, 0, 0] == pytest.approx(expect, rel=1e-6)
    assert ns == 2 * W + 1


@pytest.mark.parametrize("face", ["entry", "exit"])
def test_cone_forward_clip_boxes(face):
    """Q3: SW clips to [0,N], HW to [-0.5,N-0.5].  A unit slab on the first or
    last x layer has a tent profile; with a fine step the march converges to the
    tent's integral over the clipped interval."""
    D, H, W = SHAPE
    vol = np.zeros(SHAPE, np.float32)
    vol[:, :, 0 if face == "entry" else W - 1] = 1.0
    src = np.array([[-10.0, 4.0, 5.0]], np.float32)
    h = 1.0 / 128
    sw, _, _ = orc.cone_forward(vol, _axis_ray(0), src, (1, 1), SPACING, h, hw=False)
    hw, _, _ = orc.cone_forward(vol, _axis_ray(0), src, (1, 1), SPACING, h, hw=True)
    sx = SPACING[2]
    if face == "entry":
        assert sw[0, 0, 0] / sx == pytest.approx(0.5, abs=2 * h)     # integral_0^1 (1-x)
        assert hw[0, 0, 0] / sx == pytest.approx(0.875, abs=2 * h)   # integral_{-0.5}^1 tent
    else:
        assert sw[0, 0, 0] / sx == pytest.approx(1.0, abs=2 * h)     # whole tent inside [0,W]
        assert hw[0, 0, 0] / sx == pytest.approx(0.875, abs=2 * h)   # tent cut at W-0.5


@pytest.mark.parametrize("hw", [False, True])
def test_cone_forward_impulse_matches_map_coordinates(hw):
    """A single voxel seen along an oblique ray: fine-step march (f64 twin) equals
    the line integral of scipy's zero-padded linear interpolant (grid-constant)."""
    D, H, W = 10, 12, 14
    vol = np.zeros((D, H, W))
    vol[6, 4, 9] = 1.0
    vol[3, 7, 5] = 0.5
    d = np.array([0.9, 0.35, -0.25])
    d /= np.linalg.norm(d)
    target = np.array([9.1, 4.2, 6.1])
    s = target - 20 * d
    inv64 = np.zeros((1, 9))
    inv64[0, [2, 5, 8]] = d
    src64 = s[None]
    sp = np.array([1.0, 1.0, 1.0])
    h = 1.0 / 256
    out, _ = orc.cone_forward_f64(vol, inv64, src64, (1, 1), sp, h, hw=hw)
    t = np.linspace(0, 60, 200001)
    pts = s[None, :] + t[:, None] * d[None, :]
    f = ndimage.map_coordinates(vol, pts[:, ::-1].T, order=1, mode="grid-constant", cval=0.0)
    ref = np.trapezoid(f, t)
    assert ref > 0.3
    assert out[0, 0, 0] == pytest.approx(ref, rel=5e-3)


def test_cone_forward_gaussian_blobs_match_analytic_line_integrals():
    """End-to-end geometry pin: P matrices -> FP64 setup -> f32 march, against the
    closed-form line integral of Gaussian blobs, sigma*sqrt(2 pi)*exp(-d^2/2 sigma^2)."""
    D, H, W = 64, 56, 72
    spacing = np.array([1.0, 1.25, 0.9])            # z, y, x
    origin = geo.centred_origin((D, H, W), spacing)
    centres = np.array([[6.0, -4.0, 3.0], [-8.0, 5.0, -6.0], [1.0, 9.0, 8.0]])  # x,y,z mm
    sig = np.array([4.5, 3.75, 5.25])
    amp = np.array([1.0, 0.7, 0.5])
    zz, yy, xx = np.meshgrid(*(origin[a] + spacing[a] * np.arange(n) for a, n in enumerate((D, H, W))),
                             indexing="ij")
    vol = np.zeros((D, H, W))
    for c, sg, a in zip(centres, sig, amp):
        vol += a * np.exp(-((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2) / (2 * sg ** 2))
    Hd, Wd = 20, 28
    P = geo.circular_projection_matrices(5, 150.0, 280.0, (Hd, Wd), (2.2, 2.0), start=0.4)
    inv, src, _, _, st = orc.cone_setup(P.astype(np.float32), spacing.astype(np.float32),
                                        origin.astype(np.float32))
    assert st == 0
    out, _, _ = orc.cone_forward(vol.astype(np.float32), inv, src, (Hd, Wd), spacing, 0.25)
    ref = np.zeros_like(out, dtype=np.float64)
    vv, uu = np.meshgrid(np.arange(Hd), np.arange(Wd), indexing="ij")
    for k in range(5):
        M, p4 = P[k][:, :3], P[k][:, 3]
        S = -np.linalg.solve(M, p4)
        dirs = np.linalg.solve(M, np.stack([uu, vv, np.ones_like(uu)], 0).reshape(3, -1)).T
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        acc = np.zeros(dirs.shape[0])
        for c, sg, a in zip(centres, sig, amp):
            w = c - S
            dist2 = np.sum(w * w) - (dirs @ w) ** 2
            acc += a * sg * math.sqrt(2 * math.pi) * np.exp(-dist2 / (2 * sg ** 2))
        ref[k] = acc.reshape(Hd, Wd)
    err = np.abs(out - ref).max() / ref.max()
    assert ref.max() > 3.0
    assert err < 0.012, err   # trilinear discretisation error ~ (h/sigma)^2


def test_cone_forward_linearity_and_f64_twin():
    rng = np.random.default_rng(3)
    D, H, W = 16, 18, 20
    a = rng.uniform(0, 1, (D, H, W)).astype(np.float32)
    b = rng.uniform(0, 1, (D, H, W)).astype(np.float32)
    sp = np.array([1.0, 1.0, 1.0], np.float32)
    org = geo.centred_origin((D, H, W), sp).astype(np.float32)
    P = geo.circular_projection_matrices(4, 60.0, 100.0, (12, 16), (2.0, 2.0)).astype(np.float32)
    inv, src, inv64, src64, _ = orc.cone_setup(P, sp, org)
    pa, _, _ = orc.cone_forward(a, inv, src, (12, 16), sp, 0.5)
    pb, _, _ = orc.cone_forward(b, inv, src, (12, 16), sp, 0.5)
    pab, _, _ = orc.cone_forward(2 * a + 3 * b, inv, src, (12, 16), sp, 0.5)
    np.testing.assert_allclose(pab, 2 * pa + 3 * pb, rtol=1e-4, atol=1e-4 * pab.max())
    p64, _ = orc.cone_forward_f64(a, inv.astype(np.float64), src.astype(np.float64), (12, 16), sp, 0.5)
    rel = np.linalg.norm(pa - p64) / np.linalg.norm(p64)
    assert rel < 2e-6


@pytest.mark.parametrize("hw", [False, True])
def test_cone_forward_quantised_weights_within_bound(hw):
    rng = np.random.default_rng(5)
    D, H, W = 14, 12, 16
    vol = rng.uniform(0, 1, (D, H, W)).astype(np.float32)
    sp = np.array([1.0, 1.0, 1.0], np.float32)
    org = geo.centred_origin((D, H, W), sp).astype(np.float32)
    P = geo.circular_projection_matrices(3, 50.0, 90.0, (10, 12), (2.0, 2.0)).astype(np.float32)
    inv, src, _, _, _ = orc.cone_setup(P, sp, org)
    ex, _, _ = orc.cone_forward(vol, inv, src, (10, 12), sp, 0.5, hw=hw)
    q, bnd, _ = orc.cone_forward(vol, inv, src, (10, 12), sp, 0.5, hw=hw, qbits=8, want_bound=True)
    diff = np.abs(q - ex)
    assert diff.max() > 0
    assert np.all(diff <= 3 * 2.0 ** -9 * bnd + 1e-5)


# ----------------------------------------------------------------------------- cone backprojection (A8-A10)

@pytest.mark.parametrize("hw", [False, True])
def test_cone_backprojection_linear_sinogram(hw):
    """g_n(u,v) = a u + b v + c is reproduced exactly by bilinear interpolation, so
    each voxel receives mult * sum_n g_n(u_n, v_n) / w_n^2 with (u_n, v_n, w_n) from P."""
    D, H, W = 10, 12, 14
    sp = np.array([1.1, 0.9, 1.0])
    org = geo.centred_origin((D, H, W), sp) + np.array([0.5, -0.3, 0.2])
    Hd, Wd = 64, 72
    P = geo.circular_projection_matrices(6, 100.0, 200.0, (Hd, Wd), (1.0, 1.0), start=0.2)
    a, b, c = 0.03, -0.02, 1.5
    vv, uu = np.meshgrid(np.arange(Hd), np.arange(Wd), indexing="ij")
    sino = np.broadcast_to(a * uu + b * vv + c, (6, Hd, Wd)).astype(np.float32)
    out, _ = orc.cone_backward(sino, P.astype(np.float32), (D, H, W), sp.astype(np.float32),
                               org.astype(np.float32), mult=0.75, hw=hw)
    zz, yy, xx = np.meshgrid(*(org[a_] + sp[a_] * np.arange(n) for a_, n in enumerate((D, H, W))),
                             indexing="ij")
    X = np.stack([xx, yy, zz, np.ones_like(xx)], -1)
    ref = np.zeros((D, H, W))
    for k in range(6):
        q = X @ P[k].T
        u, v, w = q[..., 0] / q[..., 2], q[..., 1] / q[..., 2], q[..., 2]
        assert u.min() > 1 and u.max() < Wd - 2 and v.min() > 1 and v.max() < Hd - 2
        ref += (a * u + b * v + c) / w ** 2
    ref *= 0.75
    np.testing.assert_allclose(out, ref, rtol=2e-5)


def test_cone_backprojection_edge_quirk():
    """Q11: with the reference's `x > W` bounds test the right neighbour of the last
    column is the first pixel of the next row; the build's zero-outside rule reads 0.
    An affine P (w = 1) puts voxel i at u = i + 0.5, v = j."""
    Hd, Wd = 4, 5
    P = np.array([[[1, 0, 0, 0.5], [0, 1, 0, 0.0], [0, 0, 0, 1.0]]], np.float32)
    sino = np.arange(Hd * Wd, dtype=np.float32).reshape(1, Hd, Wd) + 1.0
    sp = np.ones(3, np.float32)
    org = np.zeros(3, np.float32)
    ok, _ = orc.cone_backward(sino, P, (1, Hd, Wd), sp, org)
    qk, _ = orc.cone_backward(sino, P, (1, Hd, Wd), sp, org, quirk=True)
    j = 1
    assert ok[0, j, Wd - 1] == pytest.approx(0.5 * sino[0, j, Wd - 1])
    assert qk[0, j, Wd - 1] == pytest.approx(0.5 * (sino[0, j, Wd - 1] + sino[0, j + 1, 0]))
    np.testing.assert_array_equal(ok[0, :, : Wd - 1], qk[0, :, : Wd - 1])
    # interior: (u, v) = (i + 0.5, j) -> mean of two neighbours
    assert ok[0, 2, 1] == pytest.approx(0.5 * (sino[0, 2, 1] + sino[0, 2, 2]))


def test_cone_backprojection_quantised_bound_and_f64():
    rng = np.random.default_rng(9)
    sino = rng.uniform(0, 1, (5, 20, 24)).astype(np.float32)
    P = geo.circular_projection_matrices(5, 80.0, 160.0, (20, 24), (1.5, 1.5)).astype(np.float32)
    sp = np.ones(3, np.float32)
    org = geo.centred_origin((8, 10, 12), sp).astype(np.float32)
    ex, _ = orc.cone_backward(sino, P, (8, 10, 12), sp, org, hw=True)
    q, bnd = orc.cone_backward(sino, P, (8, 10, 12), sp, org, hw=True, qbits=8, want_bound=True)
    assert np.all(np.abs(q - ex) <= 2.0 ** -8 * bnd + 1e-6)
    e64 = orc.cone_backward_f64(sino, P, (8, 10, 12), sp, org, hw=True)
    assert np.linalg.norm(ex - e64) / np.linalg.norm(e64) < 1e-6


# ----------------------------------------------------------------------------- 2D forward (A11-A13)

def test_parallel_forward_ramp_exact_with_double_spacing_scale():
    """Q7 + Q8: with spacing 0.5 the clip box is [-1, W-1] in index units, the
    volume c*(x+1) is linear on it, the march integrates in mm (x s) and the
    output is scaled by |r*s| (x s again): expected s^2 * c * W^2 / 2."""
    H = W = 16
    s = 0.5
    vol = np.tile(3.0 * (np.arange(W) + 1.0), (H, 1)).astype(np.float32)
    org = geo.centred_origin((H, W), (s, s))
    det = 40
    rays = np.array([[1.0, 0.0]], np.float32)
    out, _, _ = orc.parallel_forward(vol, rays, det, org, (s, s), -(det - 1) / 2 * 0.2, 0.2, step_factor=0.2)
    u = -(det - 1) / 2 * 0.2 + 0.2 * np.arange(det)
    yidx = (u - org[0]) / s
    inside = (yidx >= 0) & (yidx <= H - 1)
    assert inside.sum() > 10
    np.testing.assert_allclose(out[0, inside], s * s * 3.0 * W * W / 2, rtol=1e-5)
    # Rays that miss the box give 0.
    assert np.all(out[0, (yidx < -1.5) | (yidx > H + 0.5)] == 0)


@pytest.mark.parametrize("fan", [False, True])
def test_2d_forward_matches_analytic_ellipses(fan):
    H = W = 128
    s = (1.0, 1.0)
    org = geo.centred_origin((H, W), s)
    e = ph.random_ellipses(7, 12, (H, W), s, axes_range_px=(8, 45))
    img = ph.rasterise_ellipses(e, (H, W), s, org)
    n, det, ds = 30, 183, 1.0
    do = -(det - 1) / 2 * ds
    rays = geo.parallel_ray_vectors(n, math.pi if not fan else 2 * math.pi)
    if fan:
        sid, sdd = 400.0, 600.0
        out, _, _ = orc.fan_forward(img, rays, det, org, s, do, ds, sid, sdd, step_factor=0.2)
    else:
        out, _, _ = orc.parallel_forward(img, rays, det, org, s, do, ds, step_factor=0.2)
    u = do + ds * np.arange(det)
    r = rays[:, None, :]
    uv = np.stack([-rays[:, 1], rays[:, 0]], 1)[:, None, :]
    if fan:
        src = -sid * r
        detp = src + sdd * r + uv * u[None, :, None]
        ref = ph.ellipse_line_integrals(e, np.broadcast_to(src, detp.shape), detp - src)
    else:
        ref = ph.ellipse_line_integrals(e, uv * u[None, :, None], np.broadcast_to(r, (n, det, 2)))
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert rel < 0.03, rel


@pytest.mark.parametrize("fan", [False, True])
def test_2d_forward_quantised_bound_and_f64(fan):
    rng = np.random.default_rng(11)
    img = rng.uniform(0, 1, (24, 28)).astype(np.float32)
    org = geo.centred_origin((24, 28), (1.0, 1.0))
    rays = geo.parallel_ray_vectors(9, 2 * math.pi).astype(np.float32)
    kw = dict(sid=80.0, sdd=120.0) if fan else {}
    f = orc.fan_forward if fan else orc.parallel_forward
    ex, _, _ = f(img, rays, 45, org, (1.0, 1.0), -22.0, 1.0, **kw)
    q, bnd, _ = f(img, rays, 45, org, (1.0, 1.0), -22.0, 1.0, qbits=8, want_bound=True, **kw)
    assert np.all(np.abs(q - ex) <= 2 * 2.0 ** -9 * bnd + 1e-5)
    e64, _, _ = f(img, rays, 45, org, (1.0, 1.0), -22.0, 1.0, f64=True, **kw)
    # f32 alpha accumulation (Q5) drifts by ~ulp(alpha) per step: a few 1e-5.
    assert np.linalg.norm(ex - e64) / np.linalg.norm(e64) < 5e-5


# ----------------------------------------------------------------------------- 2D backprojection (A14-A15)

def test_parallel_backprojection_linear_sinogram():
    """g(s_idx) = a s_idx + b along the detector: pixel x receives
    pi/n * sum_n (a s_n + b), s_n = ((x . (-r_y, r_x)) - d_o) / d_s."""
    H, W = 20, 24
    sp = (0.8, 1.1)
    org = geo.centred_origin((H, W), sp) + np.array([0.3, -0.2])
    n, det, ds = 13, 61, 0.9
    do = -(det - 1) / 2 * ds
    rays = geo.parallel_ray_vectors(n, math.pi, start=0.1)
    a, b = 0.05, 2.0
    sino = np.broadcast_to(a * np.arange(det) + b, (n, det)).astype(np.float32)
    out, _ = orc.parallel_backward(sino, rays, (H, W), org, sp, do, ds)
    X, Y = np.meshgrid(org[1] + sp[1] * np.arange(W), org[0] + sp[0] * np.arange(H), indexing="xy")
    ref = np.zeros((H, W))
    for k in range(n):
        s = X * (-rays[k, 1]) + Y * rays[k, 0]
        si = (s - do) / ds
        assert si.min() > 0 and si.max() < det - 1
        ref += a * si + b
    ref *= math.pi / n
    np.testing.assert_allclose(out, ref, rtol=2e-5)


def test_fan_backprojection_linear_sinogram_perspective():
    """Fan: the reference intersects two lines; here the detector coordinate comes
    from the perspective ratio t = SDD / ((x - S) . r), s = t (x - S) . u."""
    H, W = 18, 22
    sp = (1.0, 0.9)
    org = geo.centred_origin((H, W), sp)
    n, det, ds = 11, 81, 0.7
    do = -(det - 1) / 2 * ds
    sid, sdd = 60.0, 95.0
    rays = geo.fan_central_ray_vectors(n, 2 * math.pi, start=0.05)
    a, b = -0.02, 1.0
    sino = np.broadcast_to(a * np.arange(det) + b, (n, det)).astype(np.float32)
    out, _ = orc.fan_backward(sino, rays, (H, W), org, sp, do, ds, sid, sdd)
    X, Y = np.meshgrid(org[1] + sp[1] * np.arange(W), org[0] + sp[0] * np.arange(H), indexing="xy")
    ref = np.zeros((H, W))
    for k in range(n):
        r = rays[k]
        u = np.array([-r[1], r[0]])
        S = -sid * r
        dx, dy = X - S[0], Y - S[1]
        t = sdd / (dx * r[0] + dy * r[1])
        s = t * (dx * u[0] + dy * u[1])
        si = (s - do) / ds
        assert si.min() > 0 and si.max() < det - 1
        ref += (a * si + b) / (dx * dx + dy * dy)
    ref *= sid * sdd * math.pi / n
    np.testing.assert_allclose(out, ref, rtol=5e-5)


@pytest.mark.parametrize("fan", [False, True])
def test_2d_backprojection_quantised_bound_and_f64(fan):
    rng = np.random.default_rng(13)
    sino = rng.uniform(0, 1, (9, 41)).astype(np.float32)
    org = geo.centred_origin((16, 18), (1.0, 1.0))
    rays = geo.parallel_ray_vectors(9, 2 * math.pi).astype(np.float32)
    kw = dict(sid=70.0, sdd=110.0) if fan else {}
    f = orc.fan_backward if fan else orc.parallel_backward
    ex, _ = f(sino, rays, (16, 18), org, (1.0, 1.0), -20.0, 1.0, **kw)
    q, bnd = f(sino, rays, (16, 18), org, (1.0, 1.0), -20.0, 1.0, qbits=8, want_bound=True, **kw)
    assert np.all(np.abs(q - ex) <= 2.0 ** -9 * bnd + 1e-5)
    e64, _ = f(sino, rays, (16, 18), org, (1.0, 1.0), -20.0, 1.0, f64=True, **kw)
    assert np.linalg.norm(ex - e64) / np.linalg.norm(e64) < 1e-5


def test_invalid_step_gives_nan_not_a_hang():
    vol = np.ones((4, 4, 4), np.float32)
    out, _, ns = orc.cone_forward(vol, _axis_ray(0), np.array([[-3.0, 1, 1]], np.float32), (1, 1),
                                  np.ones(3, np.float32), 0.0)
    assert np.isnan(out[0, 0, 0]) and ns == 0


def long_row_case(extra=1000):
    """A 1 x 1 x (2^22 + extra) volume and a 3 x 3 detector whose centre ray runs exactly
    along x through the whole row: about 8.4 million samples at step 0.5, more than 2^22
    and with alpha below 2^23, so alpha += 0.5 stays exact (beyond 2^23 it would stall at
    an even value and the reference's march would never end)."""
    W = (1 << 22) + extra
    vol = np.random.default_rng(11).uniform(0.5, 1.0, (1, 1, W)).astype(np.float32)
    f = 1.0e6
    K = np.array([[f, 0, 1.0], [0, f, 1.0], [0, 0, 1.0]])
    R = np.array([[0, 1.0, 0], [0, 0, 1.0], [1.0, 0, 0]])    # camera axes: world y, z, x
    S = np.array([-100.0, 0.5, 0.5])                          # source (x, y, z), 100 voxels before the row
    P = np.zeros((1, 3, 4))
    P[0, :, :3] = K @ R
    P[0, :, 3] = -(K @ R) @ S
    return vol, P.astype(np.float32), np.ones(3, np.float32), np.zeros(3, np.float32)


def test_cone_forward_long_row_is_not_truncated():
    """The iteration cap (at most 2^25) does not cut a valid march short: the float march
    of an 8.4-million-sample ray takes as many samples as its FP64 twin, and its line
    integral agrees with the FP64 one to the accumulated float rounding."""
    vol, P, sp, og = long_row_case()
    inv, src, inv64, src64, st = orc.cone_setup(P, sp, og)
    assert st == 0
    r, _, ns = orc.cone_forward(vol, inv, src, (3, 3), sp, 0.5)
    r64, ns64 = orc.cone_forward_f64(vol, inv64, src64, (3, 3), sp, 0.5)
    assert ns == ns64 > (1 << 22) + (1 << 22)
    assert abs(r[0, 1, 1] / r64[0, 1, 1] - 1) < 1e-4

