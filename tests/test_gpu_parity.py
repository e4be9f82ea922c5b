"""GPU parity: HIP kernels (through the C ABI) against the CPU oracle.

FP32 bar (stated tolerance): relative L2 <= 1e-5 and max|diff| <= 1e-4 * max|ref|
for the same inputs; in practice every element agrees to the last bit.
BF16 / FP16 storage: the same bar against the oracle fed the dequantised
input (the arithmetic is FP32 in both), plus relative L2 <= 1e-3 against the
FP64 oracle on the FP32 input.
"""
import math

import numpy as np
import pytest

import oracle as orc
from pyronn_layers import geometry as geo
from pyronn_layers import phantoms as ph

pytestmark = pytest.mark.gpu

gh = pytest.importorskip("gpu_helpers")


@pytest.fixture(scope="module", autouse=True)
def _cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


# ----------------------------------------------------------------------------- cone fixtures

@pytest.fixture(scope="module")
def cone():
    D, H, W = 48, 44, 40
    spacing = np.array([1.2, 1.0, 0.9], np.float32)      # z, y, x
    origin = geo.centred_origin((D, H, W), spacing).astype(np.float32)
    e = ph.random_ellipsoids(3, 20, (D, H, W), spacing, axes_range_vox=(3, 18))
    vol = ph.rasterise_ellipsoids(e, (D, H, W), spacing, origin)
    Hd, Wd = 56, 64
    P = geo.circular_projection_matrices(24, 150.0, 250.0, (Hd, Wd), (1.4, 1.3), start=0.1).astype(np.float32)
    inv, src, inv64, src64, st = orc.cone_setup(P, spacing, origin)
    assert st == 0
    return dict(vol=vol, spacing=spacing, origin=origin, P=P, det=(Hd, Wd), inv=inv, src=src, inv64=inv64,
                src64=src64, shape=(D, H, W), step=0.5)


def test_geometry_setup_bit_exact(cone):
    inv, src = gh.cone_setup(cone["P"], cone["spacing"], cone["origin"])
    np.testing.assert_array_equal(inv[0], cone["inv"])
    np.testing.assert_array_equal(src[0], cone["src"])


def test_geometry_setup_singular_gives_nan():
    P = np.zeros((1, 2, 3, 4), np.float32)
    P[0, 0] = geo.circular_projection_matrices(1, 500.0, 900.0, (8, 8), (1.0, 1.0))[0]
    P[0, 1, :, :3] = [[1, 2, 3], [2, 4, 6], [0, 0, 1]]
    inv, src = gh.cone_setup(P, np.ones(3, np.float32), np.zeros(3, np.float32))
    assert np.all(np.isfinite(inv[0, 0])) and np.all(np.isnan(inv[0, 1])) and np.all(np.isnan(src[0, 1]))


@pytest.mark.parametrize("hw", [False, True])
def test_cone_forward_parity(cone, hw):
    g = gh.cone_forward(cone["vol"], cone["P"], cone["spacing"], cone["origin"], cone["det"], cone["step"], hw=hw)
    r, _, _ = orc.cone_forward(cone["vol"], cone["inv"], cone["src"], cone["det"], cone["spacing"], cone["step"],
                               hw=hw)
    l2, mx, exact = gh.parity(g[0], r)
    print(f"cone fwd hw={hw}: rel L2 {l2:.2e} max {mx:.2e} bit-exact {exact:.4f}")
    assert exact > 0.99


@pytest.mark.parametrize("hw", [False, True])
def test_cone_backward_parity(cone, hw):
    rng = np.random.default_rng(1)
    sino, _, _ = orc.cone_forward(cone["vol"], cone["inv"], cone["src"], cone["det"], cone["spacing"], 0.5)
    sino = (sino + rng.normal(0, 0.01 * sino.max(), sino.shape)).astype(np.float32)
    g = gh.cone_backward(sino, cone["P"], cone["spacing"], cone["origin"], cone["shape"], mult=0.8, hw=hw)
    r, _ = orc.cone_backward(sino, cone["P"], cone["shape"], cone["spacing"], cone["origin"], mult=0.8, hw=hw)
    l2, mx, exact = gh.parity(g[0], r)
    print(f"cone bp hw={hw}: rel L2 {l2:.2e} max {mx:.2e} bit-exact {exact:.4f}")
    assert exact > 0.99


def _ray_dirs_f32(inv, Hd, Wd):
    """The forward kernel's per-ray direction, (u, w, 1) through inv_AR with the same
    FMA nesting, in float32 (float64 holds each product exactly)."""
    m = inv.astype(np.float64).reshape(-1, 9)
    w, u = np.meshgrid(np.arange(Hd, dtype=np.float64), np.arange(Wd, dtype=np.float64), indexing="ij")
    f = lambda a, b, c: (a * b + c).astype(np.float32).astype(np.float64)
    return np.stack([f(u[None], m[:, 3 * i, None, None], f(w[None], m[:, 3 * i + 1, None, None],
                                                            m[:, 3 * i + 2, None, None])) for i in range(3)], -1)


@pytest.mark.parametrize("case", ["axis_rays", "far_source", "tiny_step"])
@pytest.mark.parametrize("hw", [False, True])
def test_cone_forward_guarded_march(case, hw):
    """The forward marches without the coordinate clamp and the iteration cap only in
    waves whose every ray provably needs neither (all direction components non-zero,
    |source| < 2^18 and |alpha| < 2^19 voxels, step >= 2^-20 |alpha|).  These geometries
    break one condition each, so the guarded march runs; it must still match the oracle
    bit for bit."""
    shape = (12, 10, 14)
    sp = np.ones(3, np.float32)
    og = geo.centred_origin(shape, sp).astype(np.float32)
    rng = np.random.default_rng(8)
    vol = rng.uniform(0.0, 1.0, shape).astype(np.float32)
    step = 0.5
    if case == "axis_rays":      # one view at angle 0, odd detector: a row and a column of rays
        det = (15, 17)           # have an exactly zero direction component
        P = geo.circular_projection_matrices(1, 512.0, 1024.0, det, (1.0, 1.0)).astype(np.float32)
    elif case == "far_source":   # source beyond 2^18 voxels
        det = (12, 14)
        P = geo.circular_projection_matrices(4, 4.0e5, 4.0e5 + 200.0, det, (1.0, 1.0), start=0.3).astype(np.float32)
    else:                        # step below 2^-20 of the ray parameter
        det = (6, 8)
        P = geo.circular_projection_matrices(2, 100.0, 200.0, det, (2.5, 2.5), start=0.2).astype(np.float32)
        step = 2.0e-5
    inv, src, _, _, _ = orc.cone_setup(P, sp, og)
    if case == "axis_rays":
        assert (_ray_dirs_f32(inv, *det) == 0).any()
    elif case == "far_source":
        assert np.abs(src).max() >= 2.0 ** 18
    g = gh.cone_forward(vol, P, sp, og, det, step, hw=hw)[0]
    r, _, ns = orc.cone_forward(vol, inv, src, det, sp, step, hw=hw)
    assert ns > 0 and np.abs(r).max() > 0
    np.testing.assert_array_equal(g, r)


@pytest.mark.parametrize("hw", [False, True])
def test_cone_forward_long_row(hw):
    """A ray of about 8.4 million samples (more than 2^22) through a 1 x 1 x (2^22 + 1000)
    volume: the guarded march (the ray is axis-parallel) runs to the end like the oracle's,
    bit for bit."""
    from test_oracle_pins import long_row_case
    vol, P, sp, og = long_row_case()
    inv, src, _, _, _ = orc.cone_setup(P, sp, og)
    g = gh.cone_forward(vol, P, sp, og, (3, 3), 0.5, hw=hw)[0]
    r, _, ns = orc.cone_forward(vol, inv, src, (3, 3), sp, 0.5, hw=hw)
    assert ns > (1 << 23)
    np.testing.assert_array_equal(g, r)



def test_cone_hw_mode_within_texture_weight_bound(cone):
    """hardware_interp against a texture unit's fixed-point filter weights: the HIP result
    (exact weights, HW clip box and sample positions) differs from the oracle with weights
    quantised to 1/256 per axis by no more than the per-element bound that oracle computes
    alongside (forward: 3 * 2^-9 * bound; backprojection: 2^-8 * bound)."""
    g = gh.cone_forward(cone["vol"], cone["P"], cone["spacing"], cone["origin"], cone["det"], cone["step"],
                        hw=True)[0]
    q, bnd, _ = orc.cone_forward(cone["vol"], cone["inv"], cone["src"], cone["det"], cone["spacing"],
                                 cone["step"], hw=True, qbits=8, want_bound=True)
    d = np.abs(g - q)
    print(f"cone fwd hw vs 8-bit weights: {int((d > 0).sum())} of {d.size} differ, max {d.max():.3e}")
    assert d.max() > 0
    assert np.all(d <= 3 * 2.0 ** -9 * bnd + 1e-5)
    rng = np.random.default_rng(4)
    sino = rng.uniform(0, 1, (cone["P"].shape[0],) + cone["det"]).astype(np.float32)
    gb = gh.cone_backward(sino, cone["P"], cone["spacing"], cone["origin"], cone["shape"], hw=True)[0]
    qb, bb = orc.cone_backward(sino, cone["P"], cone["shape"], cone["spacing"], cone["origin"], hw=True, qbits=8,
                               want_bound=True)
    db = np.abs(gb - qb)
    print(f"cone bp hw vs 8-bit weights: {int((db > 0).sum())} of {db.size} differ, max {db.max():.3e}")
    assert np.all(db <= 2.0 ** -8 * bb + 1e-6)


def test_cone_backprojection_edge_quirk_report(cone):
    """The reference's bilinear bounds test (x > W_d, y > H_d) reads the first pixel of
    the next detector row (or the next view) for samples past the last column (row);
    this library reads zero there.  With a detector smaller than the volume's shadow some
    voxels sample that edge: report how many differ from the oracle's emulation of the
    reference reads and by how much (a reported difference, not a failure), and check that
    every other voxel is identical."""
    Hd, Wd = 30, 34
    P = geo.circular_projection_matrices(12, 150.0, 250.0, (Hd, Wd), (1.4, 1.3), start=0.1).astype(np.float32)
    rng = np.random.default_rng(6)
    sino = rng.uniform(0.5, 1.5, (12, Hd, Wd)).astype(np.float32)
    g = gh.cone_backward(sino, P, cone["spacing"], cone["origin"], cone["shape"])[0]
    ok, _ = orc.cone_backward(sino, P, cone["shape"], cone["spacing"], cone["origin"])
    qk, _ = orc.cone_backward(sino, P, cone["shape"], cone["spacing"], cone["origin"], quirk=True)
    np.testing.assert_array_equal(g, ok)
    d = np.abs(g - qk)
    hit = d > 0
    print(f"edge quirk: {int(hit.sum())} of {g.size} voxels differ from the reference's reads, "
          f"max |diff| {d.max():.3e} ({d.max() / np.abs(qk).max():.2e} of max)")
    assert 0 < hit.sum() < g.size


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_cone_forward_low_precision_storage(cone, storage):
    g = gh.cone_forward(cone["vol"], cone["P"], cone["spacing"], cone["origin"], cone["det"], cone["step"],
                        storage=storage)[0]
    vq = gh.quantise(cone["vol"], storage)
    r, _, _ = orc.cone_forward(vq, cone["inv"], cone["src"], cone["det"], cone["spacing"], cone["step"])
    gh.parity(g, r)
    r64, _ = orc.cone_forward_f64(cone["vol"], cone["inv64"], cone["src64"], cone["det"], cone["spacing"],
                                  cone["step"])
    err = np.linalg.norm(g - r64) / np.linalg.norm(r64)
    print(f"cone fwd {storage}: rel L2 vs FP64 {err:.2e}")
    assert err <= 1e-3


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_cone_backward_low_precision_storage(cone, storage):
    sino, _, _ = orc.cone_forward(cone["vol"], cone["inv"], cone["src"], cone["det"], cone["spacing"], 0.5)
    g = gh.cone_backward(sino, cone["P"], cone["spacing"], cone["origin"], cone["shape"], storage=storage)[0]
    sq = gh.quantise(sino, storage)
    r, _ = orc.cone_backward(sq, cone["P"], cone["shape"], cone["spacing"], cone["origin"])
    gh.parity(g, r)
    r64 = orc.cone_backward_f64(sino, cone["P"], cone["shape"], cone["spacing"], cone["origin"])
    err = np.linalg.norm(g - r64) / np.linalg.norm(r64)
    print(f"cone bp {storage}: rel L2 vs FP64 {err:.2e}")
    assert err <= 1e-3


def test_cone_batch_items_match_single_calls(cone):
    """Per-item geometry in one launch: each item equals its own single call, bitwise."""
    vol = cone["vol"]
    P2 = geo.circular_projection_matrices(24, 170.0, 260.0, cone["det"], (1.4, 1.3), start=0.7).astype(np.float32)
    vols = np.stack([vol, vol[::-1].copy()])
    Ps = np.stack([cone["P"], P2])
    sp = np.stack([cone["spacing"], cone["spacing"] * 1.1])
    og = np.stack([cone["origin"], cone["origin"] * 1.1])
    both = gh.cone_forward(vols, Ps, sp, og, cone["det"], [0.5, 0.6])
    for b in range(2):
        one = gh.cone_forward(vols[b], Ps[b], sp[b], og[b], cone["det"], [0.5, 0.6][b])
        np.testing.assert_array_equal(both[b], one[0])
    bp = gh.cone_backward(both, Ps, sp, og, cone["shape"], mult=np.array([1.0, 2.0]))
    for b in range(2):
        one = gh.cone_backward(both[b], Ps[b], sp[b], og[b], cone["shape"], mult=[1.0, 2.0][b])
        np.testing.assert_array_equal(bp[b], one[0])


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_cone_16bit_batch_items_match_the_oracle(cone, storage):
    """16-bit storage, two items with their own volumes and geometry in one launch: each item's
    cells are addressed by byte offsets from its own biased base; both equal the oracle."""
    vol = cone["vol"]
    P2 = geo.circular_projection_matrices(24, 170.0, 260.0, cone["det"], (1.4, 1.3), start=0.7).astype(np.float32)
    vols = np.stack([vol, vol[::-1].copy()])
    Ps = np.stack([cone["P"], P2])
    sp = np.stack([cone["spacing"], cone["spacing"] * 1.1])
    og = np.stack([cone["origin"], cone["origin"] * 1.1])
    steps = [0.5, 0.6]
    both = gh.cone_forward(vols, Ps, sp, og, cone["det"], steps, storage=storage)
    for b in range(2):
        inv, src, _, _, _ = orc.cone_setup(Ps[b], sp[b], og[b])
        r, _, _ = orc.cone_forward(gh.quantise(vols[b], storage), inv, src, cone["det"], sp[b], steps[b])
        gh.parity(both[b], r)


@pytest.mark.parametrize("hw", [False, True])
def test_cone_backprojection_voxels_on_the_source_plane(hw):
    """An orbit about z has P[0, 2] = P[2, 2] = 0, so the backprojection forms one w and one
    reciprocal per lane and view for all of the lane's z voxels.  Here the source (SID 3)
    sits on a plane of voxel centres (x = 3), where w is exactly 0: those lanes take the
    per-voxel path and reproduce the reference's infinities and NaNs bit for bit, and every
    other voxel matches as well."""
    shape = (9, 9, 9)
    sp = np.ones(3, np.float32)
    og = geo.centred_origin(shape, sp).astype(np.float32)
    P = geo.circular_projection_matrices(4, 3.0, 6.0, (12, 12), (1.0, 1.0)).astype(np.float32)
    assert np.all(P[:, 0, 2] == 0) and np.all(P[:, 2, 2] == 0)
    sino = np.random.default_rng(7).uniform(0.5, 1.5, (4, 12, 12)).astype(np.float32)
    g = gh.cone_backward(sino, P, sp, og, shape, hw=hw)[0]
    r, _ = orc.cone_backward(sino, P, shape, sp, og, hw=hw)
    assert (~np.isfinite(r)).sum() > 0
    assert np.array_equal(g, r, equal_nan=True), int((~((g == r) | (np.isnan(g) & np.isnan(r)))).sum())


def test_cone_edge_cases():
    # 1x1x1 volume, 1x1 detector, a ray that misses, and a non-positive step.
    vol = np.ones((1, 1, 1), np.float32)
    sp = np.ones(3, np.float32)
    og = np.zeros(3, np.float32)
    P = geo.circular_projection_matrices(2, 10.0, 20.0, (1, 1), (1.0, 1.0)).astype(np.float32)
    inv, src, _, _, _ = orc.cone_setup(P, sp, og)
    g = gh.cone_forward(vol, P, sp, og, (1, 1), 0.25)
    r, _, _ = orc.cone_forward(vol, inv, src, (1, 1), sp, 0.25)
    np.testing.assert_allclose(g[0], r, rtol=1e-6)
    assert g.max() > 0
    # a detector pixel far off-axis misses the 1-voxel volume
    P3 = geo.circular_projection_matrices(2, 10.0, 20.0, (3, 3), (50.0, 50.0)).astype(np.float32)
    g3 = gh.cone_forward(vol, P3, sp, og, (3, 3), 0.25)
    assert g3[0, :, 0, 0].max() == 0.0
    bad = gh.cone_forward(vol, P, sp, og, (1, 1), 0.0)
    assert np.all(np.isnan(bad))


# ----------------------------------------------------------------------------- 2D

@pytest.fixture(scope="module")
def par():
    H = W = 128
    s = (1.0, 1.0)
    org = geo.centred_origin((H, W), s).astype(np.float32)
    e = ph.random_ellipses(21, 12, (H, W), s, axes_range_px=(5, 45))
    img = ph.rasterise_ellipses(e, (H, W), s, org)
    n, det, ds = 90, 183, 1.0
    return dict(img=img, org=org, s=np.array(s, np.float32), rays=geo.parallel_ray_vectors(n).astype(np.float32),
                det=det, do=np.float32(-(det - 1) / 2 * ds), ds=np.float32(ds), shape=(H, W))


@pytest.fixture(scope="module")
def fan():
    H = W = 160
    s = (0.5, 0.5)
    org = geo.centred_origin((H, W), s).astype(np.float32)
    e = ph.random_ellipses(22, 16, (H, W), s, axes_range_px=(5, 60))
    img = ph.rasterise_ellipses(e, (H, W), s, org)
    n, det, ds = 120, 241, 1.0
    return dict(img=img, org=org, s=np.array(s, np.float32),
                rays=geo.fan_central_ray_vectors(n, 2 * math.pi).astype(np.float32), det=det,
                do=np.float32(-(det - 1) / 2 * ds), ds=np.float32(ds), shape=(H, W), sid=np.float32(1000.0),
                sdd=np.float32(1200.0))


@pytest.mark.parametrize("step", [0.2, 0.5])
def test_parallel_forward_parity(par, step):
    g = gh.forward2d(False, par["img"], par["rays"], par["det"], par["org"], par["s"], par["do"], par["ds"],
                     step_factor=step)[0]
    r, _, _ = orc.parallel_forward(par["img"], par["rays"], par["det"], par["org"], par["s"], par["do"],
                                   par["ds"], step_factor=step)
    l2, mx, exact = gh.parity(g, r)
    print(f"par fwd step {step}: rel L2 {l2:.2e} max {mx:.2e} bit-exact {exact:.4f}")


def test_parallel_backward_parity(par):
    sino, _, _ = orc.parallel_forward(par["img"], par["rays"], par["det"], par["org"], par["s"], par["do"],
                                      par["ds"])
    g = gh.backward2d(False, sino, par["rays"], par["shape"], par["org"], par["s"], par["do"], par["ds"])[0]
    r, _ = orc.parallel_backward(sino, par["rays"], par["shape"], par["org"], par["s"], par["do"], par["ds"])
    l2, mx, exact = gh.parity(g, r)
    print(f"par bp: rel L2 {l2:.2e} max {mx:.2e} bit-exact {exact:.4f}")


@pytest.mark.parametrize("step", [0.2, 0.5])
def test_fan_forward_parity(fan, step):
    g = gh.forward2d(True, fan["img"], fan["rays"], fan["det"], fan["org"], fan["s"], fan["do"], fan["ds"],
                     fan["sid"], fan["sdd"], step_factor=step)[0]
    r, _, _ = orc.fan_forward(fan["img"], fan["rays"], fan["det"], fan["org"], fan["s"], fan["do"], fan["ds"],
                              fan["sid"], fan["sdd"], step_factor=step)
    l2, mx, exact = gh.parity(g, r)
    print(f"fan fwd step {step}: rel L2 {l2:.2e} max {mx:.2e} bit-exact {exact:.4f}")


def test_fan_backward_parity(fan):
    rng = np.random.default_rng(2)
    sino, _, _ = orc.fan_forward(fan["img"], fan["rays"], fan["det"], fan["org"], fan["s"], fan["do"], fan["ds"],
                                 fan["sid"], fan["sdd"], step_factor=0.5)
    sino = (sino + rng.normal(0, 0.01 * sino.max(), sino.shape)).astype(np.float32)
    g = gh.backward2d(True, sino, fan["rays"], fan["shape"], fan["org"], fan["s"], fan["do"], fan["ds"],
                      fan["sid"], fan["sdd"])[0]
    r, _ = orc.fan_backward(sino, fan["rays"], fan["shape"], fan["org"], fan["s"], fan["do"], fan["ds"],
                            fan["sid"], fan["sdd"])
    l2, mx, exact = gh.parity(g, r)
    print(f"fan bp: rel L2 {l2:.2e} max {mx:.2e} bit-exact {exact:.4f}")


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_2d_low_precision_storage(par, storage):
    g = gh.forward2d(False, par["img"], par["rays"], par["det"], par["org"], par["s"], par["do"], par["ds"],
                     storage=storage)[0]
    r, _, _ = orc.parallel_forward(gh.quantise(par["img"], storage), par["rays"], par["det"], par["org"],
                                   par["s"], par["do"], par["ds"])
    gh.parity(g, r)
    sino = r
    gb = gh.backward2d(False, sino, par["rays"], par["shape"], par["org"], par["s"], par["do"], par["ds"],
                       storage=storage)[0]
    rb, _ = orc.parallel_backward(gh.quantise(sino, storage), par["rays"], par["shape"], par["org"], par["s"],
                                  par["do"], par["ds"])
    gh.parity(gb, rb)


def test_2d_batch_items_match_single_calls(par):
    imgs = np.stack([par["img"], par["img"].T.copy()])
    rays2 = geo.parallel_ray_vectors(90, math.pi, start=0.3).astype(np.float32)
    rays = np.stack([par["rays"], rays2])
    both = gh.forward2d(False, imgs, rays, par["det"], par["org"], par["s"], par["do"], par["ds"])
    for b in range(2):
        one = gh.forward2d(False, imgs[b], rays[b], par["det"], par["org"], par["s"], par["do"], par["ds"])
        np.testing.assert_array_equal(both[b], one[0])


@pytest.mark.parametrize("fan_beam", [False, True])
def test_2d_non_power_of_two_spacings_and_mixed_batch(fan_beam):
    """The 2D kernels multiply by 1/s instead of dividing when a spacing is a power of two
    (then both give the same bits).  Spacings 0.9 x 1.1 mm and 0.75 mm detector pixels take
    the division path; a batch whose second item has power-of-two spacings mixes both
    paths.  Every item must match the oracle and its single-item call."""
    H, W = 96, 112
    spacings = [np.array([0.9, 1.1], np.float32), np.array([0.5, 0.5], np.float32)]
    dss = [np.float32(0.75), np.float32(1.0)]
    n, det = 72, 181
    rays = (geo.fan_central_ray_vectors(n, 2 * math.pi) if fan_beam else geo.parallel_ray_vectors(n)).astype(
        np.float32)
    kw = dict(sid=np.float32(600.0), sdd=np.float32(900.0)) if fan_beam else {}
    imgs, orgs, dos = [], [], []
    for k, s in enumerate(spacings):
        org = geo.centred_origin((H, W), s).astype(np.float32)
        e = ph.random_ellipses(30 + k, 10, (H, W), s, axes_range_px=(4, 40))
        imgs.append(ph.rasterise_ellipses(e, (H, W), s, org))
        orgs.append(org)
        dos.append(np.float32(-(det - 1) / 2 * dss[k]))
    fwd_o = orc.fan_forward if fan_beam else orc.parallel_forward
    bwd_o = orc.fan_backward if fan_beam else orc.parallel_backward
    okw = dict(sid=kw["sid"], sdd=kw["sdd"]) if fan_beam else {}
    both = gh.forward2d(fan_beam, np.stack(imgs), rays, det, np.stack(orgs), np.stack(spacings), np.stack(dos),
                        np.stack(dss), step_factor=0.5, **kw)
    for k in range(2):
        one = gh.forward2d(fan_beam, imgs[k], rays, det, orgs[k], spacings[k], dos[k], dss[k], step_factor=0.5,
                           **kw)[0]
        np.testing.assert_array_equal(both[k], one)
        r, _, _ = fwd_o(imgs[k], rays, det, orgs[k], spacings[k], dos[k], dss[k], step_factor=0.5, **okw)
        l2, mx, exact = gh.parity(one, r)
        assert exact == 1.0, (k, exact)
        gb = gh.backward2d(fan_beam, r, rays, (H, W), orgs[k], spacings[k], dos[k], dss[k], **kw)[0]
        rb, _ = bwd_o(r, rays, (H, W), orgs[k], spacings[k], dos[k], dss[k], **okw)
        l2, mx, exact = gh.parity(gb, rb)
        assert exact == 1.0, (k, exact)


@pytest.mark.parametrize("fan_beam", [False, True])
def test_2d_forward_one_and_four_lane_marches(fan_beam):
    """Launches with fewer than 2^17 rays spread each ray's samples over four lanes
    (march2d_group); larger launches march one lane per ray.  The same image and geometry
    through both (a batch of one, and batches of four and eight) must match the oracle bit
    for bit."""
    H, W = 40, 36
    s = np.array([0.5, 0.5], np.float32)
    org = geo.centred_origin((H, W), s).astype(np.float32)
    e = ph.random_ellipses(71, 8, (H, W), s, axes_range_px=(3, 16))
    img = ph.rasterise_ellipses(e, (H, W), s, org)
    n, det = 720, 100                                   # 72,000 rays per image
    rays = (geo.fan_central_ray_vectors(n, 2 * math.pi) if fan_beam else geo.parallel_ray_vectors(n)).astype(
        np.float32)
    ds, do = np.float32(0.35), np.float32(-(det - 1) / 2 * 0.35)
    kw = dict(sid=np.float32(60.0), sdd=np.float32(90.0)) if fan_beam else {}
    fwd_o = orc.fan_forward if fan_beam else orc.parallel_forward
    r, _, _ = fwd_o(img, rays, det, org, s, do, ds, step_factor=0.2, **kw)
    one = gh.forward2d(fan_beam, img, rays, det, org, s, do, ds, step_factor=0.2, **kw)[0]
    np.testing.assert_array_equal(one, r)
    for B in (4, 8):                                    # 288,000 and 576,000 rays: one lane per ray
        many = gh.forward2d(fan_beam, np.stack([img] * B), np.stack([rays] * B), det, np.stack([org] * B),
                            np.stack([s] * B), np.full(B, do, np.float32), np.full(B, ds, np.float32),
                            step_factor=0.2, **kw)
        for b in range(B):
            np.testing.assert_array_equal(many[b], r)


@pytest.mark.parametrize("beam", ["par", "fan"])
def test_2d_within_texture_weight_bound(beam, par, fan):
    """The reference's 2D ops sample through the texture unit (8-bit fixed-point weights);
    the HIP sampler uses exact weights.  Forward and backprojection must differ from the
    oracle with weights quantised to 1/256 by no more than the per-element bound that
    oracle computes (forward: 2 * 2^-9 * bound, backprojection: 2^-9 * bound)."""
    c = par if beam == "par" else fan
    fb = beam == "fan"
    kw = dict(sid=c["sid"], sdd=c["sdd"]) if fb else {}
    f = orc.fan_forward if fb else orc.parallel_forward
    b = orc.fan_backward if fb else orc.parallel_backward
    g = gh.forward2d(fb, c["img"], c["rays"], c["det"], c["org"], c["s"], c["do"], c["ds"], **kw)[0]
    q, bnd, _ = f(c["img"], c["rays"], c["det"], c["org"], c["s"], c["do"], c["ds"], qbits=8, want_bound=True, **kw)
    d = np.abs(g - q)
    print(f"{beam} fwd vs 8-bit weights: {int((d > 0).sum())} of {d.size} differ, max {d.max():.3e}")
    assert d.max() > 0
    assert np.all(d <= 2 * 2.0 ** -9 * bnd + 1e-5)
    sino = np.ascontiguousarray(g)
    gb = gh.backward2d(fb, sino, c["rays"], c["shape"], c["org"], c["s"], c["do"], c["ds"], **kw)[0]
    qb, bb = b(sino, c["rays"], c["shape"], c["org"], c["s"], c["do"], c["ds"], qbits=8, want_bound=True, **kw)
    d = np.abs(gb - qb)
    print(f"{beam} bp vs 8-bit weights: {int((d > 0).sum())} of {d.size} differ, max {d.max():.3e}")
    assert np.all(d <= 2.0 ** -9 * bb + 1e-5)

