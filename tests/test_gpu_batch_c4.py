"""The uneven batch at its real sizes: eight cone problems 128^3 ... 512^3 (360-720 views,
detectors 256x256 ... 1024x768, SID 600-1000 mm, each with its own seeded 20-ellipsoid
phantom), projected and backprojected in one batched launch each, FP32 and BF16 storage.

Every problem is checked against the oracle: one forward view and one backprojected z slice,
bit for bit (BF16 against the oracle fed the dequantised input), and, for BF16, the relative
L2 error against the FP64 oracle on the FP32 inputs (bar 1e-3).  The serial per-item loop
these launches replace: cc/ops/cone_projector_3D_OPKernel.cc:234-247; the kernels:
cc/kernels/cone_projector_3D_CudaKernel.cu.cc:87-221,
cc/kernels/cone_backprojector_3D_CudaKernel.cu.cc:84-118.
"""
import numpy as np
import pytest

import oracle as orc
from pyronn_layers import geometry as geo
from pyronn_layers import phantoms as ph

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SIZES = [128, 192, 256, 256, 320, 384, 448, 512]
VIEWS = [360, 400, 450, 500, 560, 620, 680, 720]


def uneven_batch():
    """The eight geometries and phantoms (volumes on the GPU, float32)."""
    sids = np.linspace(600.0, 1000.0, len(SIZES))
    geoms, vols = [], []
    for k, (N, n, sid) in enumerate(zip(SIZES, VIEWS, sids)):
        det = (max(256, N * 3 // 2), 2 * N)
        g = geo.ConeGeometry((N, N, N), (1.0, 1.0, 1.0), det, (1.2, 1.2), n, float(sid), float(sid) * 1.6)
        e = ph.random_ellipsoids(100 + k, 20, g.volume_shape, g.volume_spacing,
                                 axes_range_vox=(8.0 * N / 256, 100.0 * N / 256))
        geoms.append(g)
        vols.append(ph.rasterise_ellipsoids_torch(e, g.volume_shape, g.volume_spacing, g.volume_origin,
                                                  device="cuda"))
    return geoms, vols


@pytest.fixture(scope="module")
def batch():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyronn_layers as pl
    pl.load()
    return uneven_batch()


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_uneven_batch_forward_views_match_oracle(batch, storage):
    from pyronn_layers import batched
    geoms, vols = batch
    st = torch.float32 if storage == "f32" else torch.bfloat16
    vin = [v.to(st) for v in vols]
    print(f"C4 forward {storage}: batched launch over {len(vin)} problems", flush=True)
    sinos = batched.cone_projection3d_batched(vin, geoms)
    torch.cuda.synchronize()
    for k, (g, v, s) in enumerate(zip(geoms, vin, sinos)):
        P = g.projection_matrices().astype(np.float32)
        sp = np.float32(g.volume_spacing)
        og = g.volume_origin.astype(np.float32)
        inv, src, inv64, src64, _ = orc.cone_setup(P, sp, og)
        view = (97 * k + 13) % g.n_views
        vh = v.float().cpu().numpy()          # the stored values (dequantised for BF16)
        ref, _, _ = orc.cone_forward(vh, inv, src, g.detector_shape, sp, g.step_size, views=(view, view + 1))
        got = s[view:view + 1].cpu().numpy()
        assert np.array_equal(got, ref), f"problem {k} ({g.volume_shape[0]}^3) view {view}: rel {_rel(got, ref):.2e}"
        if storage == "bf16":
            r64, _ = orc.cone_forward_f64(vols[k].double().cpu().numpy(), inv64, src64, g.detector_shape,
                                          np.asarray(g.volume_spacing), g.step_size, views=(view, view + 1))
            err = _rel(got, r64)
            print(f"  problem {k} {g.volume_shape[0]}^3 view {view}: bf16 rel L2 vs FP64 {err:.2e}")
            assert err <= 1e-3, err


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_uneven_batch_backprojection_slices_match_oracle(batch, storage):
    from pyronn_layers import batched
    geoms, vols = batch
    st = torch.float32 if storage == "f32" else torch.bfloat16
    # the batched forward's sinograms (checked above) as the backprojection input
    sinos = batched.cone_projection3d_batched(vols, geoms)
    sin = [s.to(st) for s in sinos]
    print(f"C4 backprojection {storage}: batched launch over {len(sin)} problems", flush=True)
    recs = batched.cone_backprojection3d_batched(sin, geoms)
    torch.cuda.synchronize()
    for k, (g, s, r) in enumerate(zip(geoms, sin, recs)):
        P = g.projection_matrices().astype(np.float32)
        sp = np.float32(g.volume_spacing)
        og = g.volume_origin.astype(np.float32)
        N = g.volume_shape[0]
        z = N // 2 + 3 * k - 7
        sh = s.float().cpu().numpy()
        ref, _ = orc.cone_backward(sh, P, g.volume_shape, sp, og, mult=g.projection_multiplier, zrange=(z, z + 1))
        got = r[z:z + 1].cpu().numpy()
        assert np.array_equal(got, ref), f"problem {k} ({N}^3) slice {z}: rel {_rel(got, ref):.2e}"
        if storage == "bf16":
            r64 = orc.cone_backward_f64(sinos[k].double().cpu().numpy(), P.astype(np.float64), g.volume_shape,
                                        np.asarray(g.volume_spacing, np.float64), g.volume_origin,
                                        mult=g.projection_multiplier, zrange=(z, z + 1))
            err = _rel(got, r64)
            print(f"  problem {k} {N}^3 slice {z}: bf16 rel L2 vs FP64 {err:.2e}")
            assert err <= 1e-3, err
        del sh
