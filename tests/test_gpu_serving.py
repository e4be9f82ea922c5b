"""GPU serving: mixed FDK / FBP requests through the continuous-batching server,
with and without HIP graphs; every result must equal the same request run alone,
bit for bit."""
import math

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


def _mix(storage):
    import pyronn_layers as pl
    from pyronn_layers.fdk import ConeFDK, FanFBP
    pl.load()
    plans = {
        "fan": FanFBP(geo.Geometry2D(volume_shape=(96, 96), volume_spacing=(0.5, 0.5), detector_shape=145,
                                     detector_spacing=1.0, n_views=120, angular_range=2 * math.pi, sid=300.0,
                                     sdd=400.0), storage=storage),
        "cone_s": ConeFDK(geo.ConeGeometry((24, 24, 24), (1.0, 1.0, 1.0), (40, 40), (1.0, 1.0), 60, 150.0, 250.0),
                          storage=storage),
        "cone_m": ConeFDK(geo.ConeGeometry((40, 36, 32), (1.0, 1.0, 1.0), (56, 64), (1.2, 1.2), 90, 200.0, 330.0),
                          storage=storage),
    }
    rng = np.random.default_rng(0)
    payloads = {}
    for k, p in plans.items():
        g = p.g
        shape = (1, g.n_views) + (tuple(g.detector_shape) if isinstance(g.detector_shape, tuple)
                                  else (g.detector_shape,))
        payloads[k] = [torch.tensor(rng.uniform(0, 2, shape).astype(np.float32)).to(storage).pin_memory()
                       for _ in range(2)]
    return plans, payloads


@pytest.mark.parametrize("use_graphs", [True, False])
@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_mixed_requests_match_alone(use_graphs, storage):
    from pyronn_layers.serving import GpuExecutor, Scheduler, latency_report, poisson_arrivals
    st = torch.float32 if storage == "f32" else torch.bfloat16
    plans, payloads = _mix(st)
    reqs = poisson_arrivals(3, 400.0, 24, {"fan": 0.4, "cone_s": 0.4, "cone_m": 0.2})
    for r in reqs:
        r.payload = payloads[r.kind][r.rid % 2]
    ex = GpuExecutor(plans, slots=3, use_graphs=use_graphs, keep_outputs=True)
    ex.warm()
    res = Scheduler(ex, slots=3, hbm_budget=1 << 34).run(reqs, idle=lambda dt: None)
    rep = latency_report(res)
    print(f"serving graphs={use_graphs} {storage}: {rep['completed']} req, {rep['req_per_s']:.1f} req/s, "
          f"p50 {rep['p50_ms']:.2f} ms, p99 {rep['p99_ms']:.2f} ms")
    assert rep["completed"] == len(reqs)
    assert len({r.slot for r in res}) > 1   # work was spread over slots
    alone = {}
    for r, q in zip(res, reqs):
        key = (q.kind, q.rid % 2)
        if key not in alone:
            alone[key] = plans[q.kind](q.payload.to("cuda")).cpu()
        assert torch.equal(r.output, alone[key]), f"request {q.rid} ({q.kind}) differs from its solo run"


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_chunked_cone_requests_match_alone(storage):
    """Cone classes served with overlapped view-range uploads (GpuExecutor chunked=...):
    every result equals the same request reconstructed alone by ConeFDK.run_chunked with
    the same ranges, bit for bit, and the one-shot plan bit for bit too (the FFT filter
    treats every projection image alone, and chunked backprojection continues the
    one-pass sums exactly)."""
    from pyronn_layers.serving import GpuExecutor, Scheduler, latency_report, poisson_arrivals
    st = torch.float32 if storage == "f32" else torch.bfloat16
    plans, payloads = _mix(st)
    chunks = {"cone_s": 2, "cone_m": 3}
    reqs = poisson_arrivals(4, 400.0, 24, {"fan": 0.3, "cone_s": 0.3, "cone_m": 0.4})
    for r in reqs:
        r.payload = payloads[r.kind][r.rid % 2]
    ex = GpuExecutor(plans, slots=3, use_graphs=True, keep_outputs=True, chunked=chunks)
    ex.warm()
    res = Scheduler(ex, slots=3, hbm_budget=1 << 34).run(reqs, idle=lambda dt: None)
    rep = latency_report(res)
    print(f"serving chunked {storage}: {rep['completed']} req, p50 {rep['p50_ms']:.2f} ms, p99 {rep['p99_ms']:.2f} ms")
    assert rep["completed"] == len(reqs)
    alone = {}
    for r, q in zip(res, reqs):
        key = (q.kind, q.rid % 2)
        if key not in alone:
            p = plans[q.kind]
            if q.kind in chunks:
                g = p.g
                kmax = max(b - a for a, b in p.chunk_bounds(chunks[q.kind]))
                bufs = [torch.empty((1, kmax) + tuple(g.detector_shape), dtype=st, device="cuda") for _ in range(2)]
                out = torch.empty((1,) + tuple(g.volume_shape), device="cuda")
                p.run_chunked(q.payload, out, bufs, torch.cuda.Stream(), chunks[q.kind])
                torch.cuda.synchronize()
                assert torch.equal(p(q.payload.to("cuda")), out)
                alone[key] = out.cpu()
            else:
                alone[key] = p(q.payload.to("cuda")).cpu()
        assert torch.equal(r.output, alone[key]), f"request {q.rid} ({q.kind}) differs from its solo run"


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_projection_requests_in_the_mix(storage):
    """Projection requests (cone and fan forward projections with device-resident geometry,
    projectors.py) served beside reconstructions: every served result equals the same
    request run alone, bit for bit, and the solo projections equal the oracle (on the
    dequantised input) on checked views.  The operators served:
    cc/ops/cone_projector_3D_OPKernel.cc:185-248, cc/ops/fan_projector_2D_OPKernel.cc:66-128."""
    import oracle as orc
    from pyronn_layers.projectors import ConeProjector, FanProjector
    from pyronn_layers.serving import GpuExecutor, Scheduler, latency_report, poisson_arrivals
    st = torch.float32 if storage == "f32" else torch.bfloat16
    plans, payloads = _mix(st)
    gc = geo.ConeGeometry((28, 30, 26), (1.0, 0.9, 1.1), (36, 44), (1.1, 1.3), 48, 160.0, 260.0)
    gf = geo.Geometry2D(volume_shape=(80, 88), volume_spacing=(0.5, 0.5), detector_shape=131, detector_spacing=1.0,
                        n_views=90, angular_range=2 * math.pi, sid=300.0, sdd=400.0, step_size=0.5)
    plans["cone_proj"] = ConeProjector(gc, storage=st)
    plans["fan_proj"] = FanProjector(gf, storage=st)
    ec = ph.random_ellipsoids(21, 6, gc.volume_shape, gc.volume_spacing, axes_range_vox=(2, 10))
    vc = ph.rasterise_ellipsoids(ec, gc.volume_shape, gc.volume_spacing, gc.volume_origin)
    ef = ph.random_ellipses(22, 8, gf.volume_shape, gf.volume_spacing)
    vf = ph.rasterise_ellipses(ef, gf.volume_shape, gf.volume_spacing, gf.volume_origin)
    payloads["cone_proj"] = [torch.tensor(vc[None]).to(st).pin_memory(),
                             torch.tensor(0.5 * vc[None]).to(st).pin_memory()]
    payloads["fan_proj"] = [torch.tensor(vf[None]).to(st).pin_memory(),
                            torch.tensor(2.0 * vf[None]).to(st).pin_memory()]
    reqs = poisson_arrivals(6, 400.0, 30, {"fan": 0.2, "cone_s": 0.2, "cone_m": 0.2, "cone_proj": 0.2,
                                           "fan_proj": 0.2})
    for r in reqs:
        r.payload = payloads[r.kind][r.rid % 2]
    ex = GpuExecutor(plans, slots=3, use_graphs=True, keep_outputs=True)
    ex.warm()
    res = Scheduler(ex, slots=3, hbm_budget=1 << 34).run(reqs, idle=lambda dt: None)
    rep = latency_report(res)
    print(f"serving with projections {storage}: {rep['completed']} req, per kind {rep['per_kind']}")
    assert rep["completed"] == len(reqs)
    assert {"cone_proj", "fan_proj"} <= set(rep["per_kind"])
    alone = {}
    for r, q in zip(res, reqs):
        key = (q.kind, q.rid % 2)
        if key not in alone:
            alone[key] = plans[q.kind](q.payload.to("cuda")).cpu()
        assert torch.equal(r.output, alone[key]), f"request {q.rid} ({q.kind}) differs from its solo run"
    # the solo projections against the oracle
    P = gc.projection_matrices().astype(np.float32)
    sp = np.asarray(gc.volume_spacing, np.float32)
    inv, src, _, _, _ = orc.cone_setup(P, sp, gc.volume_origin.astype(np.float32))
    vq = payloads["cone_proj"][0][0].float().numpy()
    for v in (0, 17):
        ref, _, ns = orc.cone_forward(vq, inv, src, gc.detector_shape, sp, gc.step_size, views=(v, v + 1))
        assert ns > 0
        assert np.array_equal(alone[("cone_proj", 0)][0, v:v + 1].numpy(), ref), f"cone view {v}"
    rays = geo.fan_central_ray_vectors(gf.n_views).astype(np.float32)
    fq = payloads["fan_proj"][0][0].float().numpy()
    ref, _, _ = orc.fan_forward(fq, rays, gf.detector_shape, np.float32(gf.volume_origin),
                                np.float32(gf.volume_spacing), gf.detector_origin, gf.detector_spacing, gf.sid, gf.sdd,
                                step_factor=0.5)
    assert np.array_equal(alone[("fan_proj", 0)][0].numpy(), ref)


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_grouped_cone_requests_match_alone(storage):
    """Queued cone FDK requests started together share one batched backprojection launch
    (GpuExecutor.start_group, Scheduler group_max): each result equals the request's solo
    plan run bit for bit, and the groups did form."""
    from pyronn_layers.serving import GpuExecutor, Request, Scheduler
    st = torch.float32 if storage == "f32" else torch.bfloat16
    plans, payloads = _mix(st)
    kinds = ["cone_s", "cone_m", "fan", "cone_m", "cone_s", "cone_s", "cone_m", "fan", "cone_s"]
    reqs = [Request(i, k, 0.0, payloads[k][i % 2]) for i, k in enumerate(kinds)]
    ex = GpuExecutor(plans, slots=2, use_graphs=True, keep_outputs=True, group_kinds=("cone_s", "cone_m"))
    ex.warm()
    res = Scheduler(ex, slots=2, hbm_budget=1 << 34, group_max=4).run(reqs, idle=lambda dt: None)
    assert all(r.error is None for r in res)
    admitted = {}
    for r in res:
        admitted.setdefault((r.slot, r.admitted), []).append(r.rid)
    assert max(len(v) for v in admitted.values()) >= 2, admitted
    for r, q in zip(res, reqs):
        solo = plans[q.kind](q.payload.to("cuda")).cpu()
        assert torch.equal(r.output, solo), f"request {q.rid} ({q.kind}) differs from its solo run"
