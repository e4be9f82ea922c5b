"""Parity at the benchmark configurations' real sizes.

C1 (parallel 256^2, 180 views, 363 px) and C2 (fan 512^2 @0.5 mm, 360 views, 725 px, SID 1000,
SDD 1200, noisy sinogram) are checked in full against the oracle.  C3 (cone 256^3, 360 views,
400x400) and the 512^3 / 720-view / 1024x768 headline workload are checked against the oracle
on subsets of views (forward) and z slices (backprojection), plus size-independent properties
on the whole problem: exact scaling by 2 (power-of-two scaling is exact in IEEE arithmetic, so
the output must double bit for bit), batch invariance, and BF16 error.
"""
import math

import numpy as np
import pytest

import oracle as orc
from pyronn_layers import geometry as geo
from pyronn_layers import phantoms as ph

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
gh = pytest.importorskip("gpu_helpers")


@pytest.fixture(scope="module", autouse=True)
def _cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_c1_parallel_full():
    H = W = 256
    s = (1.0, 1.0)
    org = geo.centred_origin((H, W), s)
    e = ph.random_ellipses(1, 12, (H, W), s, axes_range_px=(10, 90))
    img = ph.rasterise_ellipses(e, (H, W), s, org)
    rays = geo.parallel_ray_vectors(180, math.pi).astype(np.float32)
    det, ds = 363, 1.0
    do = -(det - 1) / 2 * ds
    for step in (0.5, 0.2):
        g = gh.forward2d(False, img, rays, det, org, s, do, ds, step_factor=step)[0]
        r, _, _ = orc.parallel_forward(img, rays, det, org, s, do, ds, step_factor=step)
        gh.parity(g, r)
    gb = gh.backward2d(False, r, rays, (H, W), org, s, do, ds)[0]
    rb, _ = orc.parallel_backward(r, rays, (H, W), org, s, do, ds)
    gh.parity(gb, rb)


def test_c2_fan_full_with_noise():
    H = W = 512
    s = (0.5, 0.5)
    org = geo.centred_origin((H, W), s)
    e = ph.random_ellipses(2, 16, (H, W), s)
    img = ph.rasterise_ellipses(e, (H, W), s, org)
    rays = geo.fan_central_ray_vectors(360, 2 * math.pi).astype(np.float32)
    det, ds = 725, 1.0
    do = -(det - 1) / 2 * ds
    g = gh.forward2d(True, img, rays, det, org, s, do, ds, 1000.0, 1200.0, step_factor=0.5)[0]
    r, _, _ = orc.fan_forward(img, rays, det, org, s, do, ds, 1000.0, 1200.0, step_factor=0.5)
    l2, mx, exact = gh.parity(g, r)
    print(f"C2 fan fwd: rel L2 {l2:.2e} bit-exact {exact:.4f}")
    rng = np.random.default_rng(2)
    noisy = (r + rng.normal(0, 0.01 * r.max(), r.shape)).astype(np.float32)
    gb = gh.backward2d(True, noisy, rays, (H, W), org, s, do, ds, 1000.0, 1200.0)[0]
    rb, _ = orc.fan_backward(noisy, rays, (H, W), org, s, do, ds, 1000.0, 1200.0)
    l2, mx, exact = gh.parity(gb, rb)
    print(f"C2 fan bp: rel L2 {l2:.2e} bit-exact {exact:.4f}")


def _cone_setup(shape, det, det_spacing, n, sid, sdd, seed, axes):
    sp = np.ones(3, np.float32)
    og = geo.centred_origin(shape, sp).astype(np.float32)
    P = geo.circular_projection_matrices(n, sid, sdd, det, det_spacing).astype(np.float32)
    e = ph.random_ellipsoids(seed, 20, shape, sp, axes_range_vox=axes)
    vol = ph.rasterise_ellipsoids_torch(e, shape, sp, og, device="cuda")
    inv, src, _, _, _ = orc.cone_setup(P, sp, og)
    return sp, og, P, vol, inv, src


def _fwd(vol_t, P, sp, og, det, step, hw=False):
    dev = vol_t.device
    return torch.ops.pyronn_layers.cone_projection3d(
        vol_t[None], [P.shape[0]] + list(det), torch.tensor(og[None], device=dev), torch.tensor(sp[None], device=dev),
        torch.tensor(P[None], device=dev), torch.full((1, 1), step, device=dev), torch.ones((1, 1), device=dev), hw)[0]


def _bp(sino_t, P, sp, og, shape, hw=False):
    dev = sino_t.device
    return torch.ops.pyronn_layers.cone_backprojection3d(
        sino_t[None], list(shape), torch.tensor(og[None], device=dev), torch.tensor(sp[None], device=dev),
        torch.tensor(P[None], device=dev), torch.full((1, 1), 0.5, device=dev), torch.ones((1, 1), device=dev),
        hw)[0]


@pytest.mark.parametrize("hw", [False, True])
def test_c3_cone_256_views_and_slices(hw):
    import pyronn_layers as pl
    pl.load()
    shape, det = (256, 256, 256), (400, 400)
    sp, og, P, vol, inv, src = _cone_setup(shape, det, (1.0, 1.0), 360, 750.0, 1200.0, 3, (8.0, 100.0))
    sino = _fwd(vol, P, sp, og, det, 0.5, hw)
    vh = vol.cpu().numpy()
    for v0 in (0, 97):
        r, _, _ = orc.cone_forward(vh, inv, src, det, sp, 0.5, hw=hw, views=(v0, v0 + 1))
        gh.parity(sino[v0:v0 + 1].cpu().numpy(), r)
    # backprojection of the whole 360-view sinogram in the same interpolation mode
    # (hardware_interp: cone_backprojector_3D_CudaKernel_hardware_interp.cu.cc:55-87)
    rec = _bp(sino, P, sp, og, shape, hw)
    sh = sino.cpu().numpy()
    for z0 in (0, 128, 254):
        rb, _ = orc.cone_backward(sh, P, shape, sp, og, hw=hw, zrange=(z0, z0 + 2))
        got = rec[z0:z0 + 2].cpu().numpy()
        gh.parity(got, rb)
        assert np.array_equal(got, rb), f"slices {z0}..{z0 + 1} hw={hw}: not bit-identical to the oracle"


@pytest.mark.parametrize("storage", [torch.bfloat16, torch.float16])
def test_c3_cone_256_16bit_forward(storage):
    """The 256^3, 360-view, 400 x 400 forward with 16-bit volume storage (the streamed march on
    32-bit byte offsets) against the oracle run on the same values widened to float32."""
    import pyronn_layers as pl
    pl.load()
    shape, det = (256, 256, 256), (400, 400)
    sp, og, P, vol, inv, src = _cone_setup(shape, det, (1.0, 1.0), 360, 750.0, 1200.0, 3, (8.0, 100.0))
    v16 = vol.to(storage)
    sino = _fwd(v16, P, sp, og, det, 0.5)
    vh = v16.float().cpu().numpy()
    for v0 in (0, 45, 263):
        r, _, _ = orc.cone_forward(vh, inv, src, det, sp, 0.5, views=(v0, v0 + 1))
        got = sino[v0:v0 + 1].cpu().numpy()
        assert np.array_equal(got, r), f"view {v0} {storage}: not bit-identical to the oracle"


def test_m1_cone_512_subsets_and_properties():
    import pyronn_layers as pl
    pl.load()
    shape, det = (512, 512, 512), (768, 1024)
    sp, og, P, vol, inv, src = _cone_setup(shape, det, (1.2, 1.2), 720, 750.0, 1200.0, 0, (16.0, 200.0))
    sino = _fwd(vol, P, sp, og, det, 0.5)
    vh = vol.cpu().numpy()
    r, _, _ = orc.cone_forward(vh, inv, src, det, sp, 0.5, views=(360, 361))
    l2, mx, exact = gh.parity(sino[360:361].cpu().numpy(), r)
    print(f"512^3 fwd view 360: rel L2 {l2:.2e} bit-exact {exact:.4f}")
    # exact power-of-two scaling of the whole problem
    sino2 = _fwd(vol * 2.0, P, sp, og, det, 0.5)
    assert torch.equal(sino2, sino * 2.0)
    rec = _bp(sino, P, sp, og, shape)
    sh = sino.cpu().numpy()
    rb, _ = orc.cone_backward(sh, P, shape, sp, og, zrange=(300, 301))
    l2, mx, exact = gh.parity(rec[300:301].cpu().numpy(), rb)
    print(f"512^3 bp slice 300: rel L2 {l2:.2e} bit-exact {exact:.4f}")
    rec2 = _bp(sino * 2.0, P, sp, og, shape)
    assert torch.equal(rec2, rec * 2.0)
    # BF16 storage on the whole problem: relative L2 against the FP32 result
    s16 = _fwd(vol.to(torch.bfloat16), P, sp, og, det, 0.5)
    err = ((s16 - sino).norm() / sino.norm()).item()
    print(f"512^3 bf16 fwd vs f32: rel L2 {err:.2e}")
    assert err < 1e-3
