"""GPU FDK / FBP plans: filter vs numpy convolution, full pipeline vs the CPU
oracle pipeline, reconstruction of a uniform object, BF16 storage error, and
HIP-graph replays against the eager result."""
import math

import numpy as np
import pytest

import oracle as orc
from pyronn_layers import geometry as geo
from pyronn_layers import phantoms as ph

from test_fdk_math import filter_rows

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyronn_layers as pl
    pl.load()   # registers torch.ops.pyronn_layers, which some tests call directly


@pytest.fixture(scope="module")
def cone_case():
    g = geo.ConeGeometry(volume_shape=(40, 36, 44), volume_spacing=(1.0, 1.1, 0.9), detector_shape=(48, 72),
                         detector_spacing=(1.2, 1.0), n_views=90, sid=150.0, sdd=260.0)
    P = g.projection_matrices()
    e = ph.random_ellipsoids(5, 6, g.volume_shape, g.volume_spacing, axes_range_vox=(3, 12))
    vol = ph.rasterise_ellipsoids(e, g.volume_shape, g.volume_spacing, g.volume_origin)
    inv, src, _, _, _ = orc.cone_setup(P.astype(np.float32), np.float32(g.volume_spacing),
                                       g.volume_origin.astype(np.float32))
    sino, _, _ = orc.cone_forward(vol, inv, src, g.detector_shape, np.float32(g.volume_spacing), 0.5)
    return g, vol, sino


def test_cone_filter_matches_numpy_convolution(cone_case):
    from pyronn_layers.fdk import ConeFDK
    g, _, sino = cone_case
    plan = ConeFDK(g)
    q = plan.filter(torch.tensor(sino[None], device="cuda"))[0].cpu().numpy()
    Hd, Wd = g.detector_shape
    dv, du = g.detector_spacing
    vv, uu = np.meshgrid(np.arange(Hd), np.arange(Wd), indexing="ij")
    cw = g.sdd / np.sqrt(g.sdd ** 2 + ((uu - (Wd - 1) / 2) * du) ** 2 + ((vv - (Hd - 1) / 2) * dv) ** 2)
    ref = filter_rows(sino.astype(np.float64) * cw[None], du * g.sid / g.sdd)
    err = np.linalg.norm(q - ref) / np.linalg.norm(ref)
    assert err < 2e-6, err


def _ramp_rows(x, tau, weights=None, colw=None):
    x = x.astype(np.float64) * (1.0 if weights is None else weights.astype(np.float64))
    y = filter_rows(x, tau)
    return y if colw is None else y * colw


@pytest.mark.parametrize("det_w", [1, 7, 16, 33, 400, 725, 1024, 4096])
@pytest.mark.parametrize("det_h", [1, 4, 5])
def test_fft_filter_matches_convolution_and_gemm(det_w, det_h):
    """The fused FFT filter (weighting, zero-padded circular convolution, column weights)
    against the float64 linear convolution and against the Toeplitz GEMM filter, for
    widths at and between the FFT sizes' limits, odd row counts (a row paired with
    zeros) and single rows."""
    from pyronn_layers import native as nat
    from pyronn_layers.fdk import _p, _stream
    rng = np.random.default_rng(det_w * 10 + det_h)
    images = 3 if det_w <= 1024 else 2
    x = rng.normal(size=(images, det_h, det_w)).astype(np.float32)
    wts = rng.uniform(0.5, 1.0, (det_h, det_w)).astype(np.float32)
    colw = rng.uniform(0.5, 1.5, det_w).astype(np.float32)
    tau = 0.85
    dev = torch.device("cuda")
    tx, tw, tc = (torch.tensor(a, device=dev) for a in (x, wts, colw))
    spec = torch.empty(nat.lib().pyronn_fdk_ramp_spectrum_size(det_w), device=dev)
    nat.call("pyronn_fdk_ramp_spectrum", det_w, tau, _p(tc), _p(spec), _stream())
    out = torch.empty_like(tx)
    nat.call("pyronn_fdk_filter_fft", _p(tx), nat.F32, images, det_h, det_w, _p(tw), _p(spec), _p(out), nat.F32,
             _stream())
    T = torch.empty((det_w, det_w), device=dev)
    nat.call("pyronn_fdk_ramp_matrix", det_w, tau, _p(tc), _p(T), _stream())
    ws_n = nat.lib().pyronn_fdk_filter_workspace_size(images, det_h, det_w)
    ws = torch.empty(ws_n, dtype=torch.uint8, device=dev)
    gemm = torch.empty_like(tx)
    nat.call("pyronn_fdk_filter", _p(tx), nat.F32, images, det_h, det_w, _p(tw), _p(T), _p(gemm), nat.F32, _p(ws),
             ws_n, _stream())
    torch.cuda.synchronize()
    ref = _ramp_rows(x, tau, wts[None], colw)
    got, gg = out.cpu().numpy(), gemm.cpu().numpy()
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    rel_gemm = np.linalg.norm(gg - ref) / np.linalg.norm(ref)
    print(f"W={det_w} H={det_h}: FFT rel L2 {rel:.2e}, GEMM rel L2 {rel_gemm:.2e}")
    assert rel < 1e-6, rel
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    # one image at a time gives the same bits: a row's result depends on its image only
    for i in range(images):
        one = torch.empty_like(tx[i:i + 1])
        nat.call("pyronn_fdk_filter_fft", _p(tx[i:i + 1].contiguous()), nat.F32, 1, det_h, det_w, _p(tw), _p(spec),
                 _p(one), nat.F32, _stream())
        torch.cuda.synchronize()
        assert torch.equal(one, out[i:i + 1])


@pytest.mark.parametrize("dt", [torch.float32, torch.bfloat16])
def test_fft_filter_vector_and_scalar_access_agree(dt):
    """Rows whose width is a multiple of 4 in aligned buffers take one vector load / store per
    four samples; the same rows at an address 1 element off take the scalar accesses.  The
    arithmetic is the same, so the bits are."""
    from pyronn_layers import native as nat
    from pyronn_layers.fdk import _DT, _p, _stream
    rng = np.random.default_rng(11)
    images, det_h, det_w = 3, 5, 256
    code = _DT[dt]
    x = torch.tensor(rng.normal(size=(images, det_h, det_w)).astype(np.float32), device="cuda").to(dt)
    wts = torch.tensor(rng.uniform(0.5, 1.0, (det_h, det_w)).astype(np.float32), device="cuda")
    spec = torch.empty(nat.lib().pyronn_fdk_ramp_spectrum_size(det_w), device="cuda")
    nat.call("pyronn_fdk_ramp_spectrum", det_w, 0.7, None, _p(spec), _stream())
    n = x.numel()
    aligned = torch.empty(n, dtype=dt, device="cuda")
    nat.call("pyronn_fdk_filter_fft", _p(x), code, images, det_h, det_w, _p(wts), _p(spec), _p(aligned), code,
             _stream())
    xs = torch.empty(n + 1, dtype=dt, device="cuda")
    xs[1:] = x.reshape(-1)
    ws = torch.empty(det_h * det_w + 1, device="cuda")
    ws[1:] = wts.reshape(-1)
    outs = torch.zeros(n + 1, dtype=dt, device="cuda")
    for pin, pw, po in ((xs[1:], wts, aligned.new_empty(n)), (x, ws[1:], aligned.new_empty(n)), (x, wts, outs[1:])):
        assert pin.data_ptr() % 16 or pw.data_ptr() % 16 or po.data_ptr() % 8
        nat.call("pyronn_fdk_filter_fft", _p(pin), code, images, det_h, det_w, _p(pw), _p(spec), _p(po), code,
                 _stream())
        torch.cuda.synchronize()
        assert torch.equal(po, aligned)


@pytest.mark.parametrize("src,dst", [(torch.bfloat16, torch.float32), (torch.float32, torch.bfloat16),
                                     (torch.float16, torch.float16), (torch.bfloat16, torch.bfloat16)])
def test_fft_filter_storage_types(src, dst):
    """16-bit inputs are read exactly (widened to f32); 16-bit outputs are the f32 result
    rounded to nearest even."""
    from pyronn_layers import native as nat
    from pyronn_layers.fdk import _DT, _p, _stream
    rng = np.random.default_rng(3)
    det_h, det_w = 6, 300
    x = torch.tensor(rng.normal(size=(2, det_h, det_w)).astype(np.float32), device="cuda").to(src)
    spec = torch.empty(nat.lib().pyronn_fdk_ramp_spectrum_size(det_w), device="cuda")
    nat.call("pyronn_fdk_ramp_spectrum", det_w, 1.0, None, _p(spec), _stream())
    f32 = torch.empty(x.shape, device="cuda")
    nat.call("pyronn_fdk_filter_fft", _p(x.float().contiguous()), nat.F32, 2, det_h, det_w, None, _p(spec), _p(f32),
             nat.F32, _stream())
    out = torch.empty(x.shape, dtype=dst, device="cuda")
    nat.call("pyronn_fdk_filter_fft", _p(x), _DT[src], 2, det_h, det_w, None, _p(spec), _p(out), _DT[dst], _stream())
    torch.cuda.synchronize()
    assert torch.equal(out, f32.to(dst))


def test_cone_fdk_matches_oracle_pipeline_and_recovers_density(cone_case):
    from pyronn_layers.fdk import ConeFDK
    g, vol, sino = cone_case
    plan = ConeFDK(g)
    rec = plan(torch.tensor(sino[None], device="cuda"))[0].cpu().numpy()
    q = plan.filter(torch.tensor(sino[None], device="cuda"))[0].cpu().numpy()
    ref, _ = orc.cone_backward(q, g.projection_matrices().astype(np.float32), g.volume_shape,
                               np.float32(g.volume_spacing), g.volume_origin.astype(np.float32),
                               mult=plan.mult_value)
    np.testing.assert_array_equal(rec, ref)   # same filtered input -> bit-exact backprojection
    # central slab of the reconstruction follows the phantom
    c = g.volume_shape[0] // 2
    sl = slice(c - 2, c + 3)
    err = np.abs(rec[sl] - vol[sl]).mean() / max(vol[sl].mean(), 1e-6)
    assert err < 0.35, err
    assert np.corrcoef(rec[sl].ravel(), vol[sl].ravel())[0, 1] > 0.9


def test_cone_fdk_bf16_storage(cone_case):
    from pyronn_layers.fdk import ConeFDK
    g, _, sino = cone_case
    f32 = ConeFDK(g)(torch.tensor(sino[None], device="cuda"))[0].cpu().numpy()
    b16 = ConeFDK(g, storage=torch.bfloat16)(torch.tensor(sino[None], device="cuda").to(torch.bfloat16))
    b16 = b16[0].cpu().numpy()
    err = np.linalg.norm(b16 - f32) / np.linalg.norm(f32)
    print(f"cone FDK bf16 vs f32 rel L2 {err:.2e}")
    assert err < 5e-3


def test_cone_fdk_graph_replays_match_eager(cone_case):
    from pyronn_layers.fdk import ConeFDK
    g, _, sino = cone_case
    plan = ConeFDK(g)
    x = torch.tensor(sino[None], device="cuda")
    eager = plan(x).clone()
    graph, sin, sout = plan.capture(x)
    rng = np.random.default_rng(0)
    for it in range(4):
        inp = x * float(1 + it) if it % 2 == 0 else torch.tensor(
            (sino * rng.uniform(0.5, 1.5)).astype(np.float32)[None], device="cuda")
        sin.copy_(inp)
        graph.replay()
        _ = plan(x * 0.5)   # eager work between replays
        torch.cuda.synchronize()
        ref = plan(inp)
        torch.cuda.synchronize()
        assert torch.equal(sout, ref), f"replay {it}"
    sin.copy_(x)
    graph.replay()
    torch.cuda.synchronize()
    assert torch.equal(sout, eager)


def test_fan_and_parallel_fbp_uniform_disc():
    from pyronn_layers.fdk import FanFBP, ParallelFBP
    H = W = 96
    e = ph.Ellipses(centres=np.zeros((1, 2)), axes=np.full((1, 2), 12.0), angle=np.zeros(1), density=np.array([1.0]))
    gf = geo.Geometry2D(volume_shape=(H, W), volume_spacing=(0.5, 0.5), detector_shape=181, detector_spacing=0.5,
                        n_views=240, angular_range=2 * math.pi, sid=80.0, sdd=130.0)
    rays = geo.fan_central_ray_vectors(gf.n_views)
    u = gf.detector_origin + gf.detector_spacing * np.arange(gf.detector_shape)
    r = rays[:, None, :]
    uv = np.stack([-rays[:, 1], rays[:, 0]], 1)[:, None, :]
    src = -gf.sid * r
    detp = src + gf.sdd * r + uv * u[None, :, None]
    sino = ph.ellipse_line_integrals(e, np.broadcast_to(src, detp.shape), detp - src).astype(np.float32)
    rec = FanFBP(gf)(torch.tensor(sino[None], device="cuda"))[0].cpu().numpy()
    c = H // 2
    assert rec[c - 8:c + 8, c - 8:c + 8].mean() == pytest.approx(1.0, rel=0.02)

    gp = geo.Geometry2D(volume_shape=(H, W), volume_spacing=(0.5, 0.5), detector_shape=141, detector_spacing=0.5,
                        n_views=180, angular_range=math.pi)
    rays = geo.parallel_ray_vectors(gp.n_views, math.pi)
    u = gp.detector_origin + gp.detector_spacing * np.arange(gp.detector_shape)
    uv = np.stack([-rays[:, 1], rays[:, 0]], 1)[:, None, :]
    sino = ph.ellipse_line_integrals(e, uv * u[None, :, None],
                                     np.broadcast_to(rays[:, None, :], (gp.n_views, gp.detector_shape, 2)))
    rec = ParallelFBP(gp)(torch.tensor(sino.astype(np.float32)[None], device="cuda"))[0].cpu().numpy()
    assert rec[c - 8:c + 8, c - 8:c + 8].mean() == pytest.approx(1.0, rel=0.02)


@pytest.mark.parametrize("storage", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("hw", [0, 1])
def test_view_range_backprojection_chains_to_the_one_pass_result(cone_case, storage, hw):
    """pyronn_cone_backprojection3d_views over consecutive view ranges (modes 0, 1, ..., 3)
    equals pyronn_cone_backprojection3d over all views bit for bit, for even and uneven
    ranges; mode 2 alone is the operator itself."""
    import ctypes
    from pyronn_layers import native as nat
    g, _, sino = cone_case
    dev = torch.device("cuda")
    n, (Hd, Wd), (D, H, W) = g.n_views, g.detector_shape, g.volume_shape
    s = torch.tensor(sino[None], device=dev).to(storage)
    P = torch.tensor(g.projection_matrices()[None], dtype=torch.float32, device=dev)
    vo = torch.tensor(np.asarray(g.volume_origin)[None], dtype=torch.float32, device=dev)
    vs = torch.tensor(np.asarray(g.volume_spacing)[None], dtype=torch.float32, device=dev)
    mu = torch.full((1, 1), 0.37, device=dev)
    ref = torch.ops.pyronn_layers.cone_backprojection3d(s, [D, H, W], vo, vs, P, torch.full((1, 1), 0.5, device=dev),
                                                        mu, bool(hw))
    code = {torch.float32: nat.F32, torch.bfloat16: nat.BF16}[storage]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    for edges in ([0, n], [0, 30, 60, 90], [0, 7, 50, 51, 90]):
        out = torch.full((1, D, H, W), float("nan"), device=dev)
        for i, (a, b) in enumerate(zip(edges[:-1], edges[1:])):
            k = b - a
            chunk = s[:, a:b].contiguous()
            ws_bytes = nat.lib().pyronn_cone_backprojection3d_workspace_size(1, k, Hd, Wd, code)
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
            mode = (1 if i > 0 else 0) | (2 if b == n else 0)
            nat.call("pyronn_cone_backprojection3d_views", p(chunk), code, 1, k, Hd, Wd, p(vo), p(vs),
                     p(P[:, a:b].contiguous()), p(mu), D, H, W, hw, mode, 0, D, p(out), p(ws), ws_bytes, stream)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), edges
    # the last range in z slabs of uneven depth, reusing its detector cells (mode bit 2)
    out = torch.full((1, D, H, W), float("nan"), device=dev)
    for i, (a, b) in enumerate(((0, 45), (45, n))):
        k = b - a
        chunk = s[:, a:b].contiguous()
        ws_bytes = nat.lib().pyronn_cone_backprojection3d_workspace_size(1, k, Hd, Wd, code)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        slabs = [(0, D)] if i == 0 else [(0, 3), (3, 17), (17, 18), (18, D)]
        for j, (z0, z1) in enumerate(slabs):
            mode = (1 if i > 0 else 0) | (2 if b == n else 0) | (4 if j > 0 else 0)
            nat.call("pyronn_cone_backprojection3d_views", p(chunk), code, 1, k, Hd, Wd, p(vo), p(vs),
                     p(P[:, a:b].contiguous()), p(mu), D, H, W, hw, mode, z0, z1, p(out), p(ws), ws_bytes, stream)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)


def test_view_range_z_slabs_with_a_batch_of_two(cone_case):
    """The view-range entry point with batch 2 (two sinograms, two geometries): z slabs of
    uneven depth, the later ones reusing the detector cells, give each item's one-pass
    backprojection bit for bit."""
    import ctypes
    from pyronn_layers import native as nat
    g, _, sino = cone_case
    dev = torch.device("cuda")
    n, (Hd, Wd), (D, H, W) = g.n_views, g.detector_shape, g.volume_shape
    g2 = geo.ConeGeometry(g.volume_shape, g.volume_spacing, g.detector_shape, g.detector_spacing, n, 170.0, 300.0)
    s = torch.stack([torch.tensor(sino, device=dev), torch.tensor(sino[::-1].copy(), device=dev) * 0.5])
    P = torch.tensor(np.stack([g.projection_matrices(), g2.projection_matrices()]), dtype=torch.float32, device=dev)
    vo = torch.tensor(np.stack([g.volume_origin, g2.volume_origin]), dtype=torch.float32, device=dev)
    vs = torch.tensor(np.stack([g.volume_spacing, g2.volume_spacing]), dtype=torch.float32, device=dev)
    mu = torch.tensor([[0.37], [1.3]], device=dev)
    ref = torch.ops.pyronn_layers.cone_backprojection3d(s, [D, H, W], vo, vs, P, torch.full((2, 1), 0.5, device=dev),
                                                        mu, False)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    ws_bytes = nat.lib().pyronn_cone_backprojection3d_workspace_size(2, n, Hd, Wd, nat.F32)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    out = torch.full((2, D, H, W), float("nan"), device=dev)
    for j, (z0, z1) in enumerate(((0, 5), (5, 6), (6, 23), (23, D))):
        nat.call("pyronn_cone_backprojection3d_views", p(s), nat.F32, 2, n, Hd, Wd, p(vo), p(vs), p(P), p(mu), D, H, W,
                 0, 2 | (4 if j else 0), z0, z1, p(out), p(ws), ws_bytes, stream)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)


def test_chunked_fdk_from_pinned_host_memory(cone_case):
    """ConeFDK.run_chunked uploads view ranges on a second stream while the first filters
    and backprojects: its result equals filtering each range and backprojecting all views
    in one pass bit for bit, and (the FFT filter treats every image alone) the one-shot
    plan too; the GEMM filter's rounding may depend on the row count of a call."""
    from pyronn_layers.fdk import ConeFDK
    g, _, sino = cone_case
    dev = torch.device("cuda")
    plan = ConeFDK(g, storage=torch.float32)
    host = torch.tensor(sino[None]).pin_memory()
    D, H, W = g.volume_shape
    ranges = plan.chunk_bounds(4)
    kmax = max(b - a for a, b in ranges)
    bufs = [torch.empty((1, kmax) + tuple(g.detector_shape), device=dev) for _ in range(2)]
    out = torch.empty((1, D, H, W), device=dev)
    plan.run_chunked(host, out, bufs, torch.cuda.Stream(), chunks=4)
    torch.cuda.synchronize()
    # the same with the last range in z slabs, each downloaded while the next is computed
    host_out = torch.full((1, D, H, W), float("nan")).pin_memory()
    out2 = torch.full((1, D, H, W), float("nan"), device=dev)
    plan.run_chunked(host, out2, bufs, torch.cuda.Stream(), chunks=4, host_out=host_out, out_slabs=3)
    torch.cuda.synchronize()
    assert torch.equal(out2, out) and torch.equal(host_out, out.cpu())
    dsino = host.to(dev)
    q = torch.cat([plan.filter(dsino[:, a:b]) for a, b in ranges], dim=1)
    ref = torch.ops.pyronn_layers.cone_backprojection3d(q, [D, H, W], plan.vo, plan.vs, plan.P, plan.step, plan.mult,
                                                        False)
    assert torch.equal(out, ref)
    whole = plan(dsino)
    assert plan.filter_impl == "fft"
    assert torch.equal(whole, out)
    plan.filter_impl = "gemm"
    whole = plan(dsino)
    rel = ((whole - out).norm() / whole.norm()).item()
    print(f"chunked FFT-filtered vs one-shot GEMM-filtered FDK: rel L2 {rel:.2e}")
    assert rel < 1e-6
