"""The FDK serving mix at its real sizes: fan 512^2 (360 views, 725 px, SID 1000, SDD 1200),
cone 256^3 (360 views, 400x400) and cone 512^3 (720 views, 1024x768, served in 8 view
ranges whose uploads overlap the reconstruction), two requests per class, FP32 and BF16
storage, through the continuous-batching Scheduler and GpuExecutor.

Each served result must equal the same request reconstructed alone, bit for bit; and the
reconstruction itself is checked against the oracle: one z slice of each cone volume (the
whole image of a fan request) equals the oracle's backprojection of the plan's filtered
projections, bit for bit (cc/kernels/cone_backprojector_3D_CudaKernel.cu.cc:84-118,
cc/kernels/fan_backprojector_2D_CudaKernel.cu.cc:29-67).
"""
import math

import numpy as np
import pytest

import oracle as orc
from pyronn_layers import geometry as geo
from pyronn_layers import phantoms as ph

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CHUNKS = {"cone512": 8}


def serving_geometries():
    return {
        "fan512": geo.Geometry2D(volume_shape=(512, 512), volume_spacing=(0.5, 0.5), detector_shape=725,
                                 detector_spacing=1.0, n_views=360, angular_range=2 * math.pi, sid=1000.0,
                                 sdd=1200.0),
        "cone256": geo.ConeGeometry((256, 256, 256), (1.0, 1.0, 1.0), (400, 400), (1.0, 1.0), 360, 750.0, 1200.0),
        "cone512": geo.ConeGeometry((512, 512, 512), (1.0, 1.0, 1.0), (768, 1024), (1.2, 1.2), 720, 750.0, 1200.0),
    }


@pytest.fixture(scope="module")
def clean_sinograms():
    """Projections (float32, on the GPU) of a seeded phantom per class."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyronn_layers as pl
    pl.load()
    ops = torch.ops.pyronn_layers
    f = lambda a: torch.tensor(np.asarray(a, np.float32), device="cuda")
    out = {}
    for k, g in serving_geometries().items():
        if k.startswith("fan"):
            e = ph.random_ellipses(11, 16, g.volume_shape, g.volume_spacing)
            img = f(ph.rasterise_ellipses(e, g.volume_shape, g.volume_spacing, g.volume_origin)[None])
            one = lambda v: f([[v]])
            out[k] = ops.fan_projection2d(img, [g.n_views, g.detector_shape], f(g.volume_origin[None]),
                                          one(g.detector_origin), f(np.asarray(g.volume_spacing)[None]),
                                          one(g.detector_spacing), one(g.sid), one(g.sdd),
                                          f(geo.fan_central_ray_vectors(g.n_views)[None]), 0.5)
        else:
            e = ph.random_ellipsoids(12, 20, g.volume_shape, g.volume_spacing,
                                     axes_range_vox=(8.0, 100.0 * g.volume_shape[0] / 256))
            vol = ph.rasterise_ellipsoids_torch(e, g.volume_shape, g.volume_spacing, g.volume_origin, device="cuda")
            out[k] = ops.cone_projection3d(vol[None], list(g.projection_shape), f(g.volume_origin[None]),
                                           f(np.asarray(g.volume_spacing)[None]), f(g.projection_matrices()[None]),
                                           f([[g.step_size]]), f([[1.0]]), False)
            del vol
    torch.cuda.synchronize()
    return out


def _solo(plan, kind, payload, st):
    """The request reconstructed alone, on its own, the way its class is served."""
    if kind in CHUNKS:
        g = plan.g
        kmax = max(b - a for a, b in plan.chunk_bounds(CHUNKS[kind]))
        bufs = [torch.empty((1, kmax) + tuple(g.detector_shape), dtype=st, device="cuda") for _ in range(2)]
        out = torch.empty((1,) + tuple(g.volume_shape), device="cuda")
        host_out = torch.empty(out.shape, dtype=torch.float32, pin_memory=True)
        plan.run_chunked(payload, out, bufs, torch.cuda.Stream(), CHUNKS[kind], host_out=host_out)
        torch.cuda.synchronize()
        return host_out.clone()
    return plan(payload.to("cuda")).cpu()


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_c5_mix_served_equals_alone_and_oracle(clean_sinograms, storage):
    from pyronn_layers.fdk import ConeFDK, FanFBP
    from pyronn_layers.serving import GpuExecutor, Request, Scheduler, latency_report
    st = torch.float32 if storage == "f32" else torch.bfloat16
    geoms = serving_geometries()
    plans = {k: (FanFBP(g, storage=st) if k.startswith("fan") else ConeFDK(g, storage=st)) for k, g in geoms.items()}
    gen = torch.Generator(device="cuda")
    payloads = {}
    for i, (k, s) in enumerate(clean_sinograms.items()):
        payloads[k] = []
        for j in range(2):
            gen.manual_seed(1000 * i + j)
            noisy = s + torch.randn(s.shape, device="cuda", generator=gen) * (0.01 * float(s.max()))
            payloads[k].append(noisy.to(st).cpu().pin_memory())
    # two requests per class, interleaved, arriving together: all slots busy at once
    kinds = ["cone512", "fan512", "cone256", "cone512", "cone256", "fan512"]
    reqs = [Request(i, k, 0.001 * i) for i, k in enumerate(kinds)]
    seen = {}
    for r in reqs:
        r.payload = payloads[r.kind][seen.setdefault(r.kind, 0)]
        seen[r.kind] += 1
    ex = GpuExecutor(plans, slots=3, use_graphs=True, keep_outputs=True, chunked=CHUNKS)
    print(f"C5 {storage}: warming 3 slots x 3 classes", flush=True)
    ex.warm()
    res = Scheduler(ex, slots=3, hbm_budget=160 << 30).run(reqs, idle=lambda dt: None)
    rep = latency_report(res)
    print(f"C5 {storage}: {rep['completed']} requests, p50 {rep['p50_ms']:.1f} ms, p99 {rep['p99_ms']:.1f} ms",
          flush=True)
    assert rep["completed"] == len(reqs)
    assert len({r.slot for r in res}) == 3
    for r, q in zip(res, reqs):
        plan = plans[q.kind]
        alone = _solo(plan, q.kind, q.payload, st)
        assert torch.equal(r.output, alone), f"request {q.rid} ({q.kind}) differs from its solo run"
        # the oracle on the plan's own filtered projections
        filt = plan.filter(q.payload.to("cuda")).float().cpu().numpy()[0]
        g = plan.g
        if q.kind.startswith("fan"):
            rays = geo.fan_central_ray_vectors(g.n_views, g.angular_range).astype(np.float32)
            ref, _ = orc.fan_backward(filt, rays, g.volume_shape, np.float32(g.volume_origin),
                                      np.float32(g.volume_spacing), g.detector_origin, g.detector_spacing, g.sid,
                                      g.sdd)
            got = r.output.numpy()[0]
        else:
            z = g.volume_shape[0] // 2 + 5 * q.rid - 11
            ref, _ = orc.cone_backward(filt, g.projection_matrices().astype(np.float32), g.volume_shape,
                                       np.float32(g.volume_spacing), g.volume_origin.astype(np.float32),
                                       mult=plan.mult_value, zrange=(z, z + 1))
            got = r.output.numpy()[0, z:z + 1]
        assert np.array_equal(got, ref), f"request {q.rid} ({q.kind}): reconstruction differs from the oracle"
        del filt
