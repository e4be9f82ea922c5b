"""Seeded random-geometry parity sweep: cone projection / backprojection with arbitrary
projection matrices (random camera rotations, principal points and focal lengths, sources
outside and inside the volume), anisotropic spacings, offset origins, odd and tiny shapes,
several step sizes, both interpolation modes and BF16/FP16 storage; 2D parallel / fan with
random view angles, spacings (power of two or not), step factors and BF16 storage.  Every output must
equal the oracle's bit for bit: the kernels and the oracle perform the same IEEE
operations in the same order, whatever the geometry."""
import math

import numpy as np
import pytest

import oracle as orc
from pyronn_layers import geometry as geo

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
gh = pytest.importorskip("gpu_helpers")


@pytest.fixture(scope="module", autouse=True)
def _cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _cone_case(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(v) for v in rng.integers(3, 33, size=3))              # (D, H, W)
    sp = rng.uniform(0.5, 2.0, size=3).astype(np.float32)                    # (z, y, x)
    og = (geo.centred_origin(shape, sp) + rng.uniform(-2, 2, size=3)).astype(np.float32)
    centre = (og + sp * (np.array(shape) - 1) / 2.0)[::-1]                   # (x, y, z) mm
    half = 0.5 * np.linalg.norm(sp * np.array(shape))
    inside = seed % 5 == 0
    Hd, Wd = (int(v) for v in rng.integers(5, 41, size=2))
    n = int(rng.integers(2, 7))
    P = np.empty((n, 3, 4))
    for k in range(n):
        R = _rotation(rng)                        # rows: detector u axis, v axis, viewing direction
        dist = (0.4 if inside else rng.uniform(1.5, 4.0)) * half
        S = centre - R[2] * dist                  # the camera looks along R[2] through the centre
        f = rng.uniform(0.8, 1.6) * (max(Hd, Wd) / 2.0) * dist / max(half, 1e-3)
        K = np.array([[f, 0, (Wd - 1) / 2 + rng.uniform(-2, 2)], [0, f * rng.uniform(0.8, 1.2), (Hd - 1) / 2],
                      [0, 0, 1.0]])
        P[k, :, :3] = K @ R
        P[k, :, 3] = -(K @ R) @ S
    step = float(rng.choice([0.25, 0.3, 0.5, 1.0]))
    vol = rng.uniform(0, 1, shape).astype(np.float32)
    vol[rng.uniform(size=shape) < 0.3] = 0.0
    return shape, sp, og, P.astype(np.float32), (Hd, Wd), step, vol


@pytest.mark.parametrize("seed", range(24))
def test_cone_random_geometry_bit_exact(seed):
    shape, sp, og, P, det, step, vol = _cone_case(seed)
    inv, src, _, _, _ = orc.cone_setup(P, sp, og)
    for hw in (False, True):
        g = gh.cone_forward(vol, P, sp, og, det, step, hw=hw)[0]
        r, _, _ = orc.cone_forward(vol, inv, src, det, sp, step, hw=hw)
        assert np.array_equal(g, r), (seed, hw, shape, det, step, int((g != r).sum()))
    for storage in ("bf16", "f16"):      # 16-bit storage: the oracle on the values the kernel reads
        vq = gh.quantise(vol, storage)
        for hw in (False, True):
            g = gh.cone_forward(vol, P, sp, og, det, step, hw=hw, storage=storage)[0]
            r, _, _ = orc.cone_forward(vq, inv, src, det, sp, step, hw=hw)
            assert np.array_equal(g, r), (seed, storage, hw, int((g != r).sum()))
    rng = np.random.default_rng(1000 + seed)
    sino = rng.uniform(-1, 2, (P.shape[0],) + det).astype(np.float32)
    for hw in (False, True):
        gb = gh.cone_backward(sino, P, sp, og, shape, mult=0.7, hw=hw)[0]
        rb, _ = orc.cone_backward(sino, P, shape, sp, og, mult=0.7, hw=hw)
        assert np.array_equal(gb, rb), (seed, "bp", hw, int((gb != rb).sum()))
    sq = gh.quantise(sino, "bf16")
    gb = gh.cone_backward(sino, P, sp, og, shape, mult=0.7, storage="bf16")[0]
    rb, _ = orc.cone_backward(sq, P, shape, sp, og, mult=0.7)
    assert np.array_equal(gb, rb), (seed, "bp bf16", int((gb != rb).sum()))


@pytest.mark.parametrize("seed", range(16))
@pytest.mark.parametrize("fan_beam", [False, True])
def test_2d_random_geometry_bit_exact(seed, fan_beam):
    rng = np.random.default_rng(500 + seed)
    H, W = (int(v) for v in rng.integers(3, 65, size=2))
    s = (np.array([0.5, 1.0], np.float32) if seed % 3 == 0 else rng.uniform(0.4, 1.7, size=2).astype(np.float32))
    org = (geo.centred_origin((H, W), s) + rng.uniform(-3, 3, size=2)).astype(np.float32)
    n, det = int(rng.integers(3, 40)), int(rng.integers(5, 97))
    ang = np.sort(rng.uniform(0, 2 * math.pi, n))
    rays = np.stack([np.cos(ang), np.sin(ang)], -1).astype(np.float32)
    ds = np.float32(1.0 if seed % 4 == 0 else rng.uniform(0.3, 1.5))
    do = np.float32(-(det - 1) / 2 * ds + rng.uniform(-2, 2))
    step = float(rng.choice([0.2, 0.35, 0.5, 1.0]))
    img = rng.uniform(0, 1, (H, W)).astype(np.float32)
    kw, okw = {}, {}
    if fan_beam:
        half = 0.5 * float(np.linalg.norm(s * np.array([H, W])))
        sid = np.float32((0.6 if seed % 5 == 0 else rng.uniform(1.5, 4.0)) * half)   # some sources inside
        sdd = np.float32(sid * rng.uniform(1.2, 2.0))
        kw, okw = dict(sid=sid, sdd=sdd), dict(sid=sid, sdd=sdd)
    fwd_o = orc.fan_forward if fan_beam else orc.parallel_forward
    bwd_o = orc.fan_backward if fan_beam else orc.parallel_backward
    g = gh.forward2d(fan_beam, img, rays, det, org, s, do, ds, step_factor=step, **kw)[0]
    r, _, _ = fwd_o(img, rays, det, org, s, do, ds, step_factor=step, **okw)
    assert np.array_equal(g, r), (seed, fan_beam, "fwd", int((g != r).sum()))
    sino = rng.uniform(-1, 2, (n, det)).astype(np.float32)
    gb = gh.backward2d(fan_beam, sino, rays, (H, W), org, s, do, ds, **kw)[0]
    rb, _ = bwd_o(sino, rays, (H, W), org, s, do, ds, **okw)
    np.testing.assert_array_equal(gb, rb)
    # bf16 storage: the oracle on the values the kernels read
    g = gh.forward2d(fan_beam, img, rays, det, org, s, do, ds, step_factor=step, storage="bf16", **kw)[0]
    r, _, _ = fwd_o(gh.quantise(img, "bf16"), rays, det, org, s, do, ds, step_factor=step, **okw)
    np.testing.assert_array_equal(g, r)
    gb = gh.backward2d(fan_beam, sino, rays, (H, W), org, s, do, ds, storage="bf16", **kw)[0]
    rb, _ = bwd_o(gh.quantise(sino, "bf16"), rays, (H, W), org, s, do, ds, **okw)
    np.testing.assert_array_equal(gb, rb)
