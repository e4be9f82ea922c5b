"""Shapes at the edges of the launch and index limits, against the oracle bit for bit.

* Volumes the reference accepts (its kernels index with uint32: any volume below 2^32
  voxels, cc/kernels/cone_projector_3D_CudaKernel.cu.cc:52-59,215,
  cc/kernels/cone_backprojector_3D_CudaKernel.cu.cc:116) in shapes that need the
  forward's 64-bit cell index: a (y, z) cell plane of 2^23 cells or more, and a cell
  array of more than 2^31 cells; and a 64 x 32768 x 8 slab.
* A forward whose workgroups exceed one dispatch (2^32 work-items): 1025 views of a
  2048 x 2048 detector run as two launches.
* A backprojection whose z blocks exceed one launch's grid (65535 per launch).
* Detectors with more than 2^24 cells per image, where the backprojection forms cell
  offsets with integer multiplies instead of the exact float FMA.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle as orc
from pyronn_layers import geometry as geo
from pyronn_layers import native as nat

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
gh = pytest.importorskip("gpu_helpers")


@pytest.fixture(scope="module", autouse=True)
def _cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _rand_volume(shape, seed, dtype=torch.float32):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.rand(shape, generator=g, device="cuda", dtype=torch.float32).to(dtype)


def _fwd_device(vol_t, P, sp, og, det, step, hw, views=None):
    """pyronn_cone_projection3d through the C ABI; returns the device sinogram [n, H_d, W_d]."""
    D, H, W = vol_t.shape
    code = {torch.float32: nat.F32, torch.bfloat16: nat.BF16, torch.float16: nat.F16}[vol_t.dtype]
    n = P.shape[0]
    Pd = gh.dev(P[None])
    spd, ogd = gh.dev(sp[None]), gh.dev(og[None])
    st, mu = gh.dev([[step]]), gh.dev([[1.0]])
    out = torch.empty((1, n) + tuple(det), device="cuda")
    ws = nat.lib().pyronn_cone_projection3d_workspace_size(1, D, H, W, n, code)
    work = torch.empty(ws, dtype=torch.uint8, device="cuda")
    nat.call("pyronn_cone_projection3d", gh.p(vol_t), code, 1, D, H, W, gh.p(ogd), gh.p(spd), gh.p(Pd), gh.p(st),
             gh.p(mu), n, det[0], det[1], int(hw), gh.p(out), gh.p(work), ws, gh.stream())
    torch.cuda.synchronize()
    del work
    return out[0]


def _check_views(vol_t, P, sp, og, det, step, hw, views):
    sino = _fwd_device(vol_t, P, sp, og, det, step, hw)
    vh = vol_t.float().cpu().numpy()
    inv, src, _, _, _ = orc.cone_setup(P, sp, og)
    for v in views:
        ref, _, ns = orc.cone_forward(vh, inv, src, det, sp, step, hw=hw, views=(v, v + 1))
        got = sino[v:v + 1].cpu().numpy()
        assert ns > 0
        assert np.array_equal(got, ref), f"view {v} hw={hw}: rel L2 {np.linalg.norm(got - ref) / np.linalg.norm(ref):.2e}"
    return sino


def _orbit(shape, n, sid, sdd, det, ds):
    """Circular orbit about z; ds = (row, column) pixel spacing in mm."""
    sp = np.ones(3, np.float32)
    og = geo.centred_origin(shape, sp).astype(np.float32)
    P = geo.circular_projection_matrices(n, sid, sdd, det, ds).astype(np.float32)
    return sp, og, P


@pytest.mark.parametrize("hw", [False, True])
def test_slab_64x32768x8(hw):
    """D=64, H=32768, W=8 (16.8M voxels; refused before by the transposed copy's grid)."""
    shape = (64, 32768, 8)
    det = (64, 256)
    sp, og, P = _orbit(shape, 4, 20000.0, 30000.0, det, (2.0, 10.0))
    print(f"slab {shape} hw={hw}", flush=True)
    vol = _rand_volume(shape, 1)
    sino = _check_views(vol, P, sp, og, det, 0.5, hw, views=(0, 1))   # rays along x, then along y (65536 samples)
    rec = gh.cone_backward(sino.cpu().numpy(), P, sp, og, shape, hw=hw)[0]
    ref, _ = orc.cone_backward(sino.cpu().numpy(), P, shape, sp, og, hw=hw, zrange=(30, 34))
    assert np.array_equal(rec[30:34], ref)


@pytest.mark.parametrize("storage", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("hw", [False, True])
def test_forward_64bit_cell_index_wide_plane(storage, hw):
    """(H + 2)(D + 3) = 8.4e6 >= 2^23 cells per x plane: the 64-bit index path."""
    shape = (2048, 4096, 2)
    assert (shape[1] + 2) * (shape[0] + 3) >= 2 ** 23
    det = (128, 128)
    # 0.5 mm columns (0.31 mm at the isocentre), so that view 1, whose rays run along y,
    # still has columns whose rays cross the 2-voxel-wide x extent
    sp, og, P = _orbit(shape, 4, 5000.0, 8000.0, det, (40.0, 0.5))
    print(f"64-bit index {shape} {storage} hw={hw}", flush=True)
    vol = _rand_volume(shape, 2, storage)
    _check_views(vol, P, sp, og, det, 0.5, hw, views=(0, 1))


@pytest.mark.parametrize("storage", [torch.bfloat16, torch.float16])
def test_forward_16bit_element_index_plane_2_20(storage):
    """(H + 2)(D + 3) = 2^20 cells per x plane: too wide for the 16-bit forward's 32-bit byte
    offsets (8 per cell below 2^23), so it takes the 32-bit element index."""
    shape = (1021, 1022, 2)
    assert (shape[1] + 2) * (shape[0] + 3) == 2 ** 20
    det = (64, 64)
    sp, og, P = _orbit(shape, 4, 3000.0, 5000.0, det, (20.0, 0.5))
    print(f"16-bit element index {shape} {storage}", flush=True)
    vol = _rand_volume(shape, 5, storage)
    _check_views(vol, P, sp, og, det, 0.5, False, views=(0, 1))


@pytest.mark.parametrize("storage", [torch.bfloat16, torch.float16])
def test_forward_16bit_byte_offsets_at_their_limit(storage):
    """1000 x 1020 x 368: 3.03e9 bytes of 16-bit cells, just within the byte offsets' limit
    (offset + 0x4B000000 below 2^32); view 0's rays run along x through the whole array."""
    shape = (1000, 1020, 368)
    plane = (shape[1] + 2) * (shape[0] + 3)
    assert 8 * plane < 2 ** 23 and 0xB5000000 - 8 * plane < 8 * (shape[2] + 2) * plane <= 0xB5000000
    det = (32, 48)
    sp, og, P = _orbit(shape, 4, 3000.0, 5000.0, det, (30.0, 12.0))
    print(f"16-bit byte offsets {shape} {storage}", flush=True)
    vol = _rand_volume(shape, 6, storage)
    _check_views(vol, P, sp, og, det, 0.5, False, views=(0, 1))


def test_forward_more_than_2_31_cells():
    """32 x 8192 x 8192 (2.1e9 voxels, below the reference's 2^32): 2.35e9 cells, BF16 storage."""
    shape = (32, 8192, 8192)
    assert (shape[2] + 2) * (shape[1] + 2) * (shape[0] + 3) > 2 ** 31
    det = (48, 64)
    sp, og, P = _orbit(shape, 4, 12000.0, 16000.0, det, (1.0, 300.0))
    print(f"2^31+ cells {shape} bf16", flush=True)
    vol = _rand_volume(shape, 3, torch.bfloat16)
    _check_views(vol, P, sp, og, det, 0.5, False, views=(0, 1))


def test_forward_views_beyond_one_dispatch():
    """1025 views of 2048 x 2048 pixels: 16384 workgroups per view, 1.68e7 workgroups in all,
    more than one dispatch holds; the first and the last view are checked."""
    shape = (32, 32, 32)
    n, det = 1025, (2048, 2048)
    sp, og, P = _orbit(shape, n, 100.0, 200.0, det, (0.05, 0.05))
    print(f"{n} views of {det}", flush=True)
    vol = _rand_volume(shape, 4)
    sino = _fwd_device(vol, P, sp, og, det, 0.5, False)
    vh = vol.cpu().numpy()
    inv, src, _, _, _ = orc.cone_setup(P, sp, og)
    for v in (0, 1023, 1024):
        ref, _, _ = orc.cone_forward(vh, inv, src, det, sp, 0.5, views=(v, v + 1))
        assert np.array_equal(sino[v:v + 1].cpu().numpy(), ref), f"view {v}"
    del sino
    torch.cuda.empty_cache()


def test_batched_forward_items_beyond_one_dispatch():
    """Two problems of 2048 x 2048 pixels, 600 and 500 views, in one batched call with one
    workgroup per item: 16384 items per view, 1.8e7 in all, more than one dispatch holds, so
    the work list runs as two launches.  The split falls in the second problem's
    fourteenth band of tile rows (a problem's items run band by band, each band over its tile
    columns and, inside each, over the views), so every view of that problem has tiles in
    both launches; whole views of both problems are checked."""
    from pyronn_layers import batched
    specs = [((32, 32, 32), 600, 7), ((20, 28, 24), 500, 8)]
    geoms, vols = [], []
    for shape, n, seed in specs:
        geoms.append(geo.ConeGeometry(shape, (1.0, 1.0, 1.0), (2048, 2048), (0.05, 0.05), n, 100.0, 200.0,
                                      step_size=0.5))
        vols.append(_rand_volume(shape, seed))
    items = sum(8 * ((2048 // 16 + 7) // 8) * (2048 // 16) * g.n_views for g in geoms)
    assert items > (1 << 24) - 1
    print(f"batched forward, {items} items", flush=True)
    sinos = batched.cone_projection3d_batched(vols, geoms, False, blocks_per_cu=0, forward_tile=0)
    torch.cuda.synchronize()
    for k, (g, v) in enumerate(zip(geoms, vols)):
        P = g.projection_matrices().astype(np.float32)
        sp = np.asarray(g.volume_spacing, np.float32)
        og = np.asarray(g.volume_origin, np.float32)
        inv, src, _, _, _ = orc.cone_setup(P, sp, og)
        vh = v.cpu().numpy()
        for view in (0, g.n_views // 2, g.n_views - 1):
            ref, _, ns = orc.cone_forward(vh, inv, src, g.detector_shape, sp, 0.5, views=(view, view + 1))
            assert ns > 0
            assert np.array_equal(sinos[k][view:view + 1].cpu().numpy(), ref), f"problem {k} view {view}"
    del sinos
    torch.cuda.empty_cache()


@pytest.mark.parametrize("batch", [1, 2])
def test_backprojection_z_blocks_beyond_one_launch(batch):
    """600000 (batch 1) or 300000 (batch 2) slices: more than 65535 z blocks x items, two
    launches; slices on both sides of the split are checked."""
    D = 600000 // batch
    shape = (D, 2, 2)
    sp = np.ones(3, np.float32)
    split = (65535 // batch) * 8              # first slice of the second launch (8 slices per f32 lane)
    og = np.array([-float(split), -0.5, -0.5], np.float32)   # slice `split` sits on the orbit plane
    det, n = (32, 16), 3
    P = geo.circular_projection_matrices(n, 40.0, 80.0, det, (0.5, 0.5)).astype(np.float32)
    rng = np.random.default_rng(5)
    sino = rng.uniform(0.0, 1.0, (batch, n) + det).astype(np.float32)
    print(f"bp z split: batch {batch}, {D} slices", flush=True)
    rec = gh.cone_backward(sino, P, sp, og, shape)
    for b in range(batch):
        ref, _ = orc.cone_backward(sino[b], P, shape, sp, og, zrange=(split - 9, split + 9))
        assert np.abs(ref).max() > 0
        assert np.array_equal(rec[b, split - 9:split + 9], ref), f"item {b}"


@pytest.mark.parametrize("storage", ["f32", "bf16"])
@pytest.mark.parametrize("hw", [False, True])
def test_backprojection_detector_beyond_2_24_cells(storage, hw):
    """A 4100 x 4100 detector ((H_d + 2)(W_d + 2) > 2^24 cells): integer cell offsets.  The
    volume's shadow lies near the far corner and partly off the detector."""
    shape = (16, 16, 16)
    det = (4100, 4100)
    sp = np.ones(3, np.float32)
    og = geo.centred_origin(shape, sp).astype(np.float32)
    P = geo.circular_projection_matrices(2, 100.0, 200.0, det, (0.01, 0.01),
                                         principal_point=(4000.0, 4050.0)).astype(np.float32)
    rng = np.random.default_rng(6)
    sino = rng.uniform(0.0, 1.0, (2,) + det).astype(np.float32)
    sq = gh.quantise(sino, storage)
    print(f"detector {det} {storage} hw={hw}", flush=True)
    rec = gh.cone_backward(sino, P, sp, og, shape, hw=hw, storage=storage)[0]
    ref, _ = orc.cone_backward(sq, P, shape, sp, og, hw=hw)
    assert np.abs(ref).max() > 0
    assert np.array_equal(rec, ref)
