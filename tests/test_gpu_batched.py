"""Uneven batches in one persistent launch == single-problem launches, bit for bit;
the autotuner only picks among result-preserving parameters."""
import numpy as np
import pytest

from pyronn_layers import geometry as geo
from pyronn_layers import phantoms as ph

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _problems():
    specs = [((16, 18, 20), (24, 28), 20, 80.0, 140.0), ((28, 24, 20), (36, 30), 33, 120.0, 190.0),
             ((32, 32, 32), (40, 48), 45, 150.0, 240.0), ((12, 30, 22), (20, 44), 17, 90.0, 200.0)]
    geoms, vols = [], []
    for k, (shape, det, n, sid, sdd) in enumerate(specs):
        g = geo.ConeGeometry(shape, (1.0, 1.1, 0.9), det, (1.3, 1.2), n, sid, sdd, step_size=0.5 + 0.05 * k,
                             projection_multiplier=1.0 + k)
        e = ph.random_ellipsoids(10 + k, 5, shape, g.volume_spacing, axes_range_vox=(2, 8))
        geoms.append(g)
        vols.append(ph.rasterise_ellipsoids(e, shape, g.volume_spacing, g.volume_origin))
    return geoms, vols


def _single_fwd(v, g, hw):
    dev = v.device
    return torch.ops.pyronn_layers.cone_projection3d(
        v[None], list(g.projection_shape), torch.tensor(g.volume_origin[None], dtype=torch.float32, device=dev),
        torch.tensor(np.asarray(g.volume_spacing)[None], dtype=torch.float32, device=dev),
        torch.tensor(g.projection_matrices()[None], dtype=torch.float32, device=dev),
        torch.full((1, 1), g.step_size, device=dev), torch.full((1, 1), g.projection_multiplier, device=dev), hw)[0]


def _single_bp(s, g, hw):
    dev = s.device
    return torch.ops.pyronn_layers.cone_backprojection3d(
        s[None], list(g.volume_shape), torch.tensor(g.volume_origin[None], dtype=torch.float32, device=dev),
        torch.tensor(np.asarray(g.volume_spacing)[None], dtype=torch.float32, device=dev),
        torch.tensor(g.projection_matrices()[None], dtype=torch.float32, device=dev),
        torch.full((1, 1), g.step_size, device=dev), torch.full((1, 1), g.projection_multiplier, device=dev), hw)[0]


@pytest.mark.parametrize("storage", [torch.float32, torch.bfloat16, torch.float16])
@pytest.mark.parametrize("hw", [False, True])
def test_batched_equals_single(storage, hw):
    import pyronn_layers as pl
    from pyronn_layers import batched
    pl.load()
    geoms, vols = _problems()
    vt = [torch.tensor(v, device="cuda").to(storage) for v in vols]
    singles = [_single_fwd(v, g, hw) for v, g in zip(vt, geoms)]
    # one workgroup per item (blocks_per_cu 0) and persistent grids, every forward tile shape
    for bpc, tile in ((0, 0), (0, 1), (0, 2), (0, 3), (-1, 0), (2, 0), (4, 1), (8, 2), (16, 3)):
        sinos = batched.cone_projection3d_batched(vt, geoms, hw, blocks_per_cu=bpc, forward_tile=tile)
        for s, s1 in zip(sinos, singles):
            assert torch.equal(s, s1), f"blocks_per_cu {bpc} tile {tile}"
    st = [s.to(storage) for s in sinos]
    for zpl, bpc in ((0, 0), (4, -1), (4, 2), (8, 16)):   # one workgroup per item, then persistent grids
        recs = batched.cone_backprojection3d_batched(st, geoms, hw, z_per_lane=zpl, blocks_per_cu=bpc)
        for r, s, g in zip(recs, st, geoms):
            assert torch.equal(r, _single_bp(s, g, hw))
    # geometry arrays kept resident on the device across calls
    dg = batched.device_geometry(geoms, "cuda")
    for s, s1 in zip(batched.cone_projection3d_batched(vt, geoms, hw, device_geometry=dg), singles):
        assert torch.equal(s, s1)
    for r, r1 in zip(batched.cone_backprojection3d_batched(st, geoms, hw, device_geometry=dg), recs):
        assert torch.equal(r, r1)


def test_autotuner_caches_choice(tmp_path):
    import pyronn_layers as pl
    from pyronn_layers.autotune import Autotuner
    pl.load()
    geoms, vols = _problems()
    vt = [torch.tensor(v, device="cuda") for v in vols]
    at = Autotuner(str(tmp_path / "tune.json"), reps=1)
    f = at.forward(vt, geoms)
    assert f["blocks_per_cu"] in (0, 8, 16)
    assert f["forward_tile"] in (0, 1, 2, 3) and "x" in f["tile_shape"]
    assert len(f["trials_ms"]) == 6                      # four tiles, then two more grid sizes
    from pyronn_layers import batched
    sinos = batched.cone_projection3d_batched(vt, geoms, blocks_per_cu=f["blocks_per_cu"],
                                              forward_tile=f["forward_tile"])
    b = at.backward(sinos, geoms)
    assert b["z_per_lane"] in (4, 8)
    again = Autotuner(str(tmp_path / "tune.json"))
    assert again.forward(vt, geoms) == f
