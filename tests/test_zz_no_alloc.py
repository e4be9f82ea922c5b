"""pyronn_cone_projection3d_noalloc (the forward that allocates nothing) against
pyronn_cone_projection3d (the workspace entry point), bit for bit, through the C ABI.

Two calls in a row share the library's fixed geometry scratch: the second, smaller call
must not see anything of the first.  Each case runs a larger problem, then a smaller one,
and compares both with the workspace entry point's outputs; one case is also checked
against the oracle.  The workspace entry point follows
cc/kernels/cone_projector_3D_CudaKernel.cu.cc:87-221 (hardware_interp:
cc/kernels/cone_projector_3D_CudaKernel_hardware_interp.cu.cc:43-182).
"""
import numpy as np
import pytest

import oracle as orc
from pyronn_layers import geometry as geo
from pyronn_layers import native as nat
from pyronn_layers import phantoms as ph

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
gh = pytest.importorskip("gpu_helpers")


@pytest.fixture(scope="module", autouse=True)
def _cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _problem(shape, n, det, seed, batch=1):
    D, H, W = shape
    sp = np.array([1.0, 0.9, 1.1], np.float32)
    og = geo.centred_origin(shape, sp).astype(np.float32)
    vols = []
    for b in range(batch):
        e = ph.random_ellipsoids(seed + b, 6, shape, sp, axes_range_vox=(2, max(3, min(shape) // 2)))
        vols.append(ph.rasterise_ellipsoids(e, shape, sp, og))
    P = np.stack([geo.circular_projection_matrices(n, 150.0, 240.0, det, (1.1, 1.3), start=0.3 * b)
                  for b in range(batch)]).astype(np.float32)
    return np.stack(vols), P, sp, og


def _noalloc(vol, P, sp, og, det, step, hw, storage):
    B, D, H, W = vol.shape
    n = P.shape[1]
    code, tdt = gh.DT[storage]
    v = gh.dev(vol, tdt)
    Pd = gh.dev(P)
    spd = gh.dev(np.broadcast_to(sp, (B, 3)))
    ogd = gh.dev(np.broadcast_to(og, (B, 3)))
    st = gh.dev(np.full((B, 1), step, np.float32))
    mu = gh.dev(np.ones((B, 1), np.float32))
    out = torch.full((B, n) + tuple(det), float("nan"), device="cuda")
    nat.call("pyronn_cone_projection3d_noalloc", gh.p(v), code, B, D, H, W, gh.p(ogd), gh.p(spd), gh.p(Pd),
             gh.p(st), gh.p(mu), n, det[0], det[1], int(hw), gh.p(out), gh.stream())
    torch.cuda.synchronize()
    return out.cpu().numpy()


CASES = [
    # (storage, hw, batch)
    ("f32", False, 1),
    ("f32", True, 1),
    ("bf16", False, 2),
    ("f16", True, 1),
]


@pytest.mark.parametrize("storage,hw,batch", CASES)
def test_noalloc_equals_workspace_entry_point_twice(storage, hw, batch):
    big = ((40, 36, 44), 30, (32, 40))
    small = ((12, 10, 14), 6, (9, 11))
    for k, (shape, n, det) in enumerate((big, small)):
        print(f"noalloc {storage} hw={hw} batch={batch} call {k}: volume {shape}, {n} views, det {det}",
              flush=True)
        vol, P, sp, og = _problem(shape, n, det, 10 * k + 1, batch)
        got = _noalloc(vol, P, sp, og, det, 0.5, hw, storage)
        ref = gh.cone_forward(vol, P, sp, og, det, 0.5, hw=hw, storage=storage)
        assert got.shape == ref.shape
        assert np.abs(ref).max() > 0
        assert np.array_equal(got, ref), f"call {k}: {np.count_nonzero(got != ref)} of {got.size} differ"


def test_noalloc_matches_oracle_after_a_larger_call():
    """A larger call first (batch of 3, 48 views), then one volume against the oracle."""
    vol, P, sp, og = _problem((30, 28, 26), 48, (24, 30), 5, batch=3)
    _noalloc(vol, P, sp, og, (24, 30), 0.5, False, "f32")
    vol, P, sp, og = _problem((18, 16, 20), 10, (14, 16), 7)
    got = _noalloc(vol, P, sp, og, (14, 16), 0.5, False, "f32")[0]
    inv, src, _, _, _ = orc.cone_setup(P[0], sp, og)
    ref, _, ns = orc.cone_forward(vol[0], inv, src, (14, 16), sp, 0.5)
    assert ns > 0
    assert np.array_equal(got, ref)


def test_noalloc_rejects_more_views_than_its_scratch():
    vol, P, sp, og = _problem((8, 8, 8), 4, (4, 4), 3)
    v = gh.dev(vol)
    Pd = gh.dev(P)
    spd, ogd = gh.dev(sp[None]), gh.dev(og[None])
    st, mu = gh.dev([[0.5]]), gh.dev([[1.0]])
    out = torch.empty(16, device="cuda")
    rc = nat.lib().pyronn_cone_projection3d_noalloc(gh.p(v), nat.F32, 2, 8, 8, 8, gh.p(ogd), gh.p(spd), gh.p(Pd),
                                                    gh.p(st), gh.p(mu), 40000, 4, 4, 0, gh.p(out), gh.stream())
    assert rc == 3 and "scratch" in nat.last_error()
