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
