"""The torch op library: reference op surface (names, input order, shape rule),
geometry broadcasting, storage dtypes, and gradient pairing (each projection's
gradient is the paired backprojection of the incoming gradient with the same
geometry inputs, and vice versa) - checked bit for bit against explicit calls."""
import math

import numpy as np
import pytest

from pyronn_layers import geometry as geo

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyronn_layers as pl
    pl.load()


def _cone_args(B, n=12, det=(20, 24), shape=(16, 18, 20), dev="cuda"):
    sp = torch.tensor([[1.0, 1.1, 0.9]] * B, device=dev)
    og = torch.tensor(geo.centred_origin(shape, (1.0, 1.1, 0.9))[None].repeat(B, 0), dtype=torch.float32, device=dev)
    P = torch.tensor(np.stack([geo.circular_projection_matrices(n, 80.0 + 5 * b, 140.0, det, (1.3, 1.3))
                               for b in range(B)]), dtype=torch.float32, device=dev)
    return og, sp, P, torch.full((B, 1), 0.5, device=dev), torch.full((B, 1), 1.5, device=dev)


@pytest.mark.parametrize("hw", [False, True])
def test_cone_shapes_and_gradient_pairing(hw):
    import pyronn_layers as pl
    B, n, det, shape = 2, 12, (20, 24), (16, 18, 20)
    og, sp, P, st, mu = _cone_args(B, n, det, shape)
    vol = torch.rand((B,) + shape, device="cuda", requires_grad=True)
    sino = pl.cone_projection3d(vol, torch.tensor([n, *det], dtype=torch.int32), og, sp, P, st, mu,
                                hardware_interp=hw)
    assert sino.shape == (B, n) + det and sino.dtype == torch.float32
    g = torch.rand_like(sino)
    sino.backward(g)
    expect = torch.ops.pyronn_layers.cone_backprojection3d(g, list(shape), og, sp, P, st, mu, hw)
    assert torch.equal(vol.grad, expect)

    s = torch.rand((B, n) + det, device="cuda", requires_grad=True)
    rec = pl.cone_backprojection3d(s, list(shape), og, sp, P, st, mu, hardware_interp=hw)
    assert rec.shape == (B,) + shape
    gv = torch.rand_like(rec)
    rec.backward(gv)
    assert torch.equal(s.grad, torch.ops.pyronn_layers.cone_projection3d(gv, [n, *det], og, sp, P, st, mu, hw))


def test_cone_geometry_broadcast_and_cpu_geometry():
    import pyronn_layers as pl
    n, det, shape = 10, (16, 18), (12, 12, 12)
    og, sp, P, st, mu = _cone_args(1, n, det, shape)
    vol = torch.rand((2,) + shape, device="cuda")
    full = pl.cone_projection3d(vol, [n, *det], og.repeat(2, 1), sp.repeat(2, 1), P.repeat(2, 1, 1, 1),
                                st.repeat(2, 1), mu.repeat(2, 1))
    bcast = pl.cone_projection3d(vol, [n, *det], og[0], sp[0], P[0], st[0], mu[0])
    cpu_geo = pl.cone_projection3d(vol, [n, *det], og.cpu(), sp.cpu(), P.cpu(), st.cpu(), mu.cpu())
    assert torch.equal(full, bcast) and torch.equal(full, cpu_geo)
    with pytest.raises(Exception):
        pl.cone_projection3d(vol.cpu(), [n, *det], og, sp, P, st, mu)
    with pytest.raises(Exception):
        pl.cone_projection3d(vol, [n, *det], og, sp, P[:, :5], st, mu)


@pytest.mark.parametrize("storage", [torch.bfloat16, torch.float16])
def test_cone_low_precision_inputs_and_grad_dtype(storage):
    import pyronn_layers as pl
    n, det, shape = 10, (16, 18), (12, 12, 12)
    og, sp, P, st, mu = _cone_args(1, n, det, shape)
    vol = torch.rand((1,) + shape, device="cuda").to(storage).requires_grad_(True)
    sino = pl.cone_projection3d(vol, [n, *det], og, sp, P, st, mu)
    assert sino.dtype == torch.float32
    ref = pl.cone_projection3d(vol.detach().float(), [n, *det], og, sp, P, st, mu)
    assert torch.equal(sino, ref)   # same values after dequantisation -> same bits
    sino.sum().backward()
    assert vol.grad.dtype == storage


def _par_args(B, n, det):
    rays = torch.tensor(np.stack([geo.parallel_ray_vectors(n, math.pi, start=0.1 * b) for b in range(B)]),
                        dtype=torch.float32, device="cuda")
    og = torch.tensor([[-15.5, -17.5]] * B, device="cuda")
    sp = torch.ones((B, 2), device="cuda")
    do = torch.full((B, 1), -(det - 1) / 2.0, device="cuda")
    ds = torch.ones((B, 1), device="cuda")
    return og, do, sp, ds, rays


def test_parallel_shapes_and_gradient_pairing():
    import pyronn_layers as pl
    B, n, det = 2, 30, 49
    og, do, sp, ds, rays = _par_args(B, n, det)
    img = torch.rand((B, 32, 36), device="cuda", requires_grad=True)
    sino = pl.parallel_projection2d(img, torch.tensor([n, det], dtype=torch.int32), og, do, sp, ds, rays)
    assert sino.shape == (B, n, det)
    g = torch.rand_like(sino)
    sino.backward(g)
    assert torch.equal(img.grad, torch.ops.pyronn_layers.parallel_backprojection2d(g, [32, 36], og, do, sp, ds, rays))
    s = torch.rand((B, n, det), device="cuda", requires_grad=True)
    rec = pl.parallel_backprojection2d(s, [32, 36], og, do, sp, ds, rays)
    gv = torch.rand_like(rec)
    rec.backward(gv)
    assert torch.equal(s.grad, torch.ops.pyronn_layers.parallel_projection2d(gv, [n, det], og, do, sp, ds, rays, 0.2))


def test_fan_shapes_and_gradient_pairing():
    import pyronn_layers as pl
    B, n, det = 2, 30, 61
    og, do, sp, ds, rays = _par_args(B, n, det)
    sid = torch.full((B, 1), 90.0, device="cuda")
    sdd = torch.full((B, 1), 150.0, device="cuda")
    img = torch.rand((B, 32, 36), device="cuda", requires_grad=True)
    sino = pl.fan_projection2d(img, [n, det], og, do, sp, ds, sid, sdd, rays, step_size=0.5)
    assert sino.shape == (B, n, det)
    g = torch.rand_like(sino)
    sino.backward(g)
    assert torch.equal(img.grad, torch.ops.pyronn_layers.fan_backprojection2d(g, [32, 36], og, do, sp, ds, sid, sdd,
                                                                               rays))
    s = torch.rand((B, n, det), device="cuda", requires_grad=True)
    rec = pl.fan_backprojection2d(s, [32, 36], og, do, sp, ds, sid, sdd, rays)
    gv = torch.rand_like(rec)
    rec.backward(gv)
    assert torch.equal(s.grad, torch.ops.pyronn_layers.fan_projection2d(gv, [n, det], og, do, sp, ds, sid, sdd,
                                                                        rays, 0.2))


def test_gradcheck_style_linearity_of_pair():
    """<A x, y> and <x, B y> are both linear in x and y: a finite check that autograd
    routes through the paired op (unmatched pair, so no adjoint identity is asserted)."""
    import pyronn_layers as pl
    n, det, shape = 8, (12, 14), (10, 10, 10)
    og, sp, P, st, mu = _cone_args(1, n, det, shape)
    x = torch.rand((1,) + shape, device="cuda", requires_grad=True)
    y = torch.rand((1, n) + det, device="cuda")
    (pl.cone_projection3d(x, [n, *det], og, sp, P, st, mu) * y).sum().backward()
    g1 = x.grad.clone()
    x.grad = None
    (pl.cone_projection3d(x, [n, *det], og, sp, P, st, mu) * (2 * y)).sum().backward()
    assert torch.equal(x.grad, 2 * g1)


def test_empty_batches_give_empty_outputs():
    """An empty batch returns an empty output of the right shape (as the reference op does)
    for all six ops, and an empty batch still validates its geometry."""
    import pyronn_layers as pl
    og, sp, P, st, mu = _cone_args(1, 6, (8, 10), (6, 6, 6))
    vol = torch.zeros((0, 6, 6, 6), device="cuda")
    assert pl.cone_projection3d(vol, [6, 8, 10], og[0], sp[0], P[0], st[0], mu[0]).shape == (0, 6, 8, 10)
    sino = torch.zeros((0, 6, 8, 10), device="cuda")
    assert pl.cone_backprojection3d(sino, [6, 6, 6], og[0], sp[0], P[0], st[0], mu[0]).shape == (0, 6, 6, 6)
    with pytest.raises(Exception):
        pl.cone_projection3d(vol, [6, 8, 10], og[0], sp[0], P[0, :3], st[0], mu[0])
    og2, do, sp2, ds, rays = _par_args(1, 12, 17)
    img = torch.zeros((0, 16, 16), device="cuda")
    assert pl.parallel_projection2d(img, [12, 17], og2[0], do[0], sp2[0], ds[0], rays[0]).shape == (0, 12, 17)
    s2 = torch.zeros((0, 12, 17), device="cuda")
    assert pl.parallel_backprojection2d(s2, [16, 16], og2[0], do[0], sp2[0], ds[0], rays[0]).shape == (0, 16, 16)
    sid = torch.full((1,), 90.0, device="cuda")
    sdd = torch.full((1,), 150.0, device="cuda")
    assert pl.fan_projection2d(img, [12, 17], og2[0], do[0], sp2[0], ds[0], sid, sdd, rays[0]).shape == (0, 12, 17)
    assert pl.fan_backprojection2d(s2, [16, 16], og2[0], do[0], sp2[0], ds[0], sid, sdd, rays[0]).shape == (0, 16, 16)
