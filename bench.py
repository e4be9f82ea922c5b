"""Throughput of the cone-beam projector / backprojector pair on one MI355X.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3]

Workload (the headline configuration): cone-beam circular orbit, volume 512^3
float32 at 1 mm with a centred origin, 720 views over 2 pi, 1024 x 768 detector
at 1.2 mm, SID 750 mm, SDD 1200 mm, step 0.5 voxel, software interpolation,
phantom of 20 seeded random ellipsoids (semi-axes U[16,200] voxels, density
U[0,1]).  One step = forward projection of the volume + backprojection of the
resulting sinogram, inputs resident in HBM.

value = (S_fwd + U_bp) / (t_fwd + t_bp) / 1e9 in Gsamples/s, where S_fwd is the
number of forward samples (entrance + loop + exit positions of every ray) and
U_bp = voxels x views.  The roofline object reports the dominant kernel: its
algorithmic bytes (gather-footprint model: 8 neighbours x input bytes per
forward sample + 4 B per output pixel; 4 neighbours x input bytes per
backprojection update + 4 B per voxel) over its average time from HIP events
on the launch stream.  A BF16-storage run (FP32 accumulation, FP32 outputs) is
reported beside it, and the CPU oracle is timed on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "pyro-nn-layers_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "cone 512³ fwd+back Gsamples/s & HBM %roofline; FDK req/s + p99, BF16 vs FP32"
HBM_PEAK = 8.0e12


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def m1_config():
    return dict(shape=(512, 512, 512), spacing=(1.0, 1.0, 1.0), n=720, det=(768, 1024), det_spacing=(1.2, 1.2),
                sid=750.0, sdd=1200.0, step=0.5, seed=0, n_ellipsoids=20, axes=(16.0, 200.0))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-bf16", action="store_true", help="skip the BF16-storage run")
    ap.add_argument("--no-hw", action="store_true", help="skip the hardware_interp run")
    ap.add_argument("--no-f16", action="store_true", help="skip the FP16-storage run")
    ap.add_argument("--cpu-views", type=int, default=12, help="forward views in the CPU sample")
    ap.add_argument("--cpu-slices", type=int, default=48, help="z slices in the CPU backprojection sample")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one plain 1 GiB device copy (reference for HBM byte counters)")
    ap.add_argument("--bf16-only", action="store_true", help="time only the BF16-storage run")
    ap.add_argument("--no-serve", action="store_true", help="skip the FDK serving run")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-rate", type=float, default=8.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-slots", type=int, default=3)
    ap.add_argument("--serve-chunks", type=int, default=8,
                    help="view ranges of a 512^3 request uploaded while earlier ranges reconstruct (1 = off)")
    ap.add_argument("--serve-chunks-256", type=int, default=1,
                    help="the same for the cone 256^3 requests (1 = off: one upload, then one HIP graph)")
    ap.add_argument("--serve-curve", type=lambda t: [float(x) for x in t.split(",") if x], default=[16.0, 32.0, 48.0, 64.0],
                    help="offered loads (req/s) of the serving latency curve, comma separated")
    ap.add_argument("--no-serve-projections", dest="serve_projections", action="store_false",
                    help="skip the serving run that mixes projection requests in")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the uneven-batch run")
    ap.add_argument("--batched-only", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the C1-C3 configuration measurements")
    ap.add_argument("--configs-only", action="store_true")
    ap.add_argument("--filter-only", action="store_true",
                    help="time the FDK ramp filter (FFT and GEMM) on one 512^3 request's sinogram")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the library has no multi-GPU path.",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import pyronn_layers as pl
    from pyronn_layers import geometry as geo
    from pyronn_layers import metrics
    from pyronn_layers import phantoms as ph

    pl.load()
    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        a = torch.ones(2 ** 28, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    # the single-section modes need neither the 512^3 phantom nor its sample count
    if args.serve_only:
        print(json.dumps(serve_bench(args)), flush=True)
        return 0
    if args.batched_only:
        print(json.dumps(batched_bench(args)), flush=True)
        return 0
    if args.configs_only:
        print(json.dumps(configs_bench(args)), flush=True)
        return 0
    if args.filter_only:
        print(json.dumps(filter_bench(args)), flush=True)
        return 0
    cfg = m1_config()
    D, H, W = cfg["shape"]
    n = cfg["n"]
    Hd, Wd = cfg["det"]
    sp = np.array(cfg["spacing"], np.float32)
    og = geo.centred_origin(cfg["shape"], sp).astype(np.float32)
    P = geo.circular_projection_matrices(n, cfg["sid"], cfg["sdd"], cfg["det"], cfg["det_spacing"]).astype(np.float32)

    log("phantom")
    e = ph.random_ellipsoids(cfg["seed"], cfg["n_ellipsoids"], cfg["shape"], sp, axes_range_vox=cfg["axes"])
    vol = ph.rasterise_ellipsoids_torch(e, cfg["shape"], sp, og, device=dev)[None].contiguous()
    vo = torch.tensor(og[None], device=dev)
    vs = torch.tensor(sp[None], device=dev)
    Pt = torch.tensor(P[None], device=dev)
    st = torch.full((1, 1), cfg["step"], device=dev)
    mu = torch.ones((1, 1), device=dev)
    ops = torch.ops.pyronn_layers

    log("sample count")
    S_fwd = metrics.cone_forward_sample_count(P, cfg["shape"], sp, og, cfg["det"], cfg["step"], device=dev)
    U_bp = metrics.cone_backprojection_update_count(cfg["shape"], n)
    R = n * Hd * Wd
    V = D * H * W

    medians = {}

    def run(storage, hw=False, keep_rec=False):
        v = vol.to(storage)
        times_f, times_b = [], []
        sino = None
        last_rec = None
        for it in range(args.warmup + args.steps):
            timed = it >= args.warmup
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            torch.cuda.synchronize()
            e0.record()
            sino = ops.cone_projection3d(v, [n, Hd, Wd], vo, vs, Pt, st, mu, hw)
            e1.record()
            s_in = sino if storage == torch.float32 else sino.to(storage)
            e1b = torch.cuda.Event(enable_timing=True)
            e1b.record()
            rec = ops.cone_backprojection3d(s_in, [D, H, W], vo, vs, Pt, st, mu, hw)
            e2.record()
            torch.cuda.synchronize()
            tf, tb = e0.elapsed_time(e1) / 1e3, e1b.elapsed_time(e2) / 1e3
            log(f"{'step' if timed else 'warmup'} {it} {str(storage)}{' hw' if hw else ''}: "
                f"fwd {tf * 1e3:.1f} ms, bp {tb * 1e3:.1f} ms")
            if timed:
                times_f.append(tf)
                times_b.append(tb)
            if keep_rec:
                last_rec = rec
            del rec
        medians[(storage, hw)] = (float(np.median(times_f)), float(np.median(times_b)))
        if keep_rec:
            return float(np.mean(times_f)), float(np.mean(times_b)), sino, last_rec
        return float(np.mean(times_f)), float(np.mean(times_b)), sino

    # Timed region: K steps bracketed by synchronize on both sides (inside run()).
    if args.bf16_only:
        tf, tb, _ = run(torch.bfloat16)
        log(f"bf16 only: fwd {tf * 1e3:.1f} ms, bp {tb * 1e3:.1f} ms")
        return 0
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    tf, tb, sino = run(torch.float32)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall0

    gs = (S_fwd + U_bp) / (tf + tb) / 1e9
    B_fwd = S_fwd * 8 * 4 + R * 4
    B_bp = U_bp * 4 * 4 + V * 4
    dominant = "cone_fwd_kernel" if tf >= tb else "cone_bp_kernel"
    B_dom, t_dom = (B_fwd, tf) if tf >= tb else (B_bp, tb)
    traffic = None
    ceiling = None
    cpath = os.path.join(ROOT, "profiles", "cone_m1_f32_counters.json")
    if dominant == "cone_fwd_kernel" and os.path.exists(cpath):
        # Counters of the same kernel on the same workload (rocprofv3 --pmc, one pass per
        # counter group): HBM bytes per launch from FETCH_SIZE / WRITE_SIZE scaled by the
        # reading of a 1 GiB copy, and the wave load instructions the address unit issues.
        with open(cpath) as f:
            ctr = json.load(f)
        traffic = ctr["traffic_bytes"]
        # The unit that binds: the texture-address path issues at most one wave load per 16
        # cycles per CU (one cycle per group of four lanes; profiles/unit_rates_mi355x.md).
        loads_per_cu_cycle = ctr["vmem_load_instructions"] / ctr["cus"] / (t_dom * ctr["shader_clock_hz"])
        ceiling = {
            "unit": "texture-address path (wave loads per CU per cycle)",
            "achieved": round(loads_per_cu_cycle, 5), "peak": ctr["load_issue_peak_per_cu_cycle"],
            "frac": round(loads_per_cu_cycle / ctr["load_issue_peak_per_cu_cycle"], 4),
            "busy_frac": ctr["ta_busy_frac"],
            "note": "the rest of the busy time is groups of four lanes reading two 128-byte lines",
            "source": "profiles/cone_m1_f32_counters.json (SQ_INSTS_VMEM_RD, shader clock), "
                      "profiles/unit_rates_mi355x.md (peak)",
        }
    out = {
        "metric": METRIC,
        "value": round(gs, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round((tf + tb) * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 20-ellipsoid phantom, generated on device)",
        "config": {
            "workload": "cone-beam 512^3 @1mm, 720 views/2pi, 1024x768 det @1.2mm, SID 750, SDD 1200, "
                        "step 0.5 vox, software interp; forward + backprojection of its sinogram",
            "volume": [D, H, W], "views": n, "detector": [Hd, Wd], "step_size": cfg["step"],
            "hardware_interp": False, "parallelism": "none (one GPU)",
        },
        "roofline": {
            "bound": "hbm", "kernel": dominant,
            # gather-footprint model: the bytes the reference's gathers request (8 corners per
            # sample), most of them served by the L1 and L2; a ratio above 1 means cache reuse
            "model": "gather footprint (algorithmic bytes / HBM peak), not DRAM traffic",
            "achieved": round(B_dom / t_dom / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(B_dom / t_dom / HBM_PEAK, 4), "traffic": traffic,
            "traffic_source": "profiles/cone_m1_f32_counters.json (rocprofv3 FETCH_SIZE+WRITE_SIZE, copy-calibrated)",
            # measured HBM bytes over the live duration: the share of HBM bandwidth actually used
            "traffic_frac": round(traffic / t_dom / HBM_PEAK, 4) if traffic else None,
            "ceiling": ceiling,
        },
        "fwd_back": {
            "t_fwd_ms": round(tf * 1e3, 2), "t_bp_ms": round(tb * 1e3, 2),
            "t_fwd_median_ms": round(medians[(torch.float32, False)][0] * 1e3, 2),
            "t_bp_median_ms": round(medians[(torch.float32, False)][1] * 1e3, 2),
            "S_fwd": S_fwd, "U_bp": U_bp,
            "fwd_gsamples_s": round(S_fwd / tf / 1e9, 2), "bp_gupdates_s": round(U_bp / tb / 1e9, 2),
            "algorithmic_bytes": B_fwd + B_bp,
            "hbm_roofline_frac_fwd_back": round((B_fwd + B_bp) / (tf + tb) / HBM_PEAK, 4),
            "wall_s_timed_region": round(wall, 3),
        },
    }

    if not args.no_bf16:
        del sino
        torch.cuda.empty_cache()
        tf16, tb16, sino16, rec16 = run(torch.bfloat16, keep_rec=True)
        prec16 = precision_vs_fp64(vol, sino16, rec16, torch.bfloat16, P, sp, og, cfg)
        del sino16, rec16
        Bf = S_fwd * 8 * 2 + R * 4
        Bb = U_bp * 4 * 2 + V * 4
        g16 = (S_fwd + U_bp) / (tf16 + tb16) / 1e9
        out["bf16"] = {
            "value": round(g16, 3), "unit": "Gsamples/s", "speedup_vs_f32": round(g16 / gs, 3),
            "t_fwd_ms": round(tf16 * 1e3, 2), "t_bp_ms": round(tb16 * 1e3, 2),
            "hbm_roofline_frac_fwd_back": round((Bf + Bb) / (tf16 + tb16) / HBM_PEAK, 4),
            "storage": "bf16 volume and sinogram, fp32 accumulation, fp32 outputs",
            "error_vs_fp64": prec16,
        }
    else:
        del sino

    if not args.no_f16:
        torch.cuda.empty_cache()
        tfh16, tbh16, sino_f16, rec_f16 = run(torch.float16, keep_rec=True)
        prec_f16 = precision_vs_fp64(vol, sino_f16, rec_f16, torch.float16, P, sp, og, cfg)
        del sino_f16, rec_f16
        Bf = S_fwd * 8 * 2 + R * 4
        Bb = U_bp * 4 * 2 + V * 4
        gh16 = (S_fwd + U_bp) / (tfh16 + tbh16) / 1e9
        out["f16"] = {
            "value": round(gh16, 3), "unit": "Gsamples/s", "speedup_vs_f32": round(gh16 / gs, 3),
            "t_fwd_ms": round(tfh16 * 1e3, 2), "t_bp_ms": round(tbh16 * 1e3, 2),
            "hbm_roofline_frac_fwd_back": round((Bf + Bb) / (tfh16 + tbh16) / HBM_PEAK, 4),
            "storage": "fp16 volume and sinogram, fp32 accumulation, fp32 outputs",
            "error_vs_fp64": prec_f16,
        }

    if not args.no_hw:
        # hardware_interp: texture clip box [-0.5, N-0.5] and sample positions, exact weights
        torch.cuda.empty_cache()
        S_hw = metrics.cone_forward_sample_count(P, cfg["shape"], sp, og, cfg["det"], cfg["step"],
                                                 hardware_interp=True, device=dev)
        tfh, tbh, sino_h = run(torch.float32, hw=True)
        del sino_h
        out["hardware_interp"] = {
            "value": round((S_hw + U_bp) / (tfh + tbh) / 1e9, 3), "unit": "Gsamples/s",
            "t_fwd_ms": round(tfh * 1e3, 2), "t_bp_ms": round(tbh * 1e3, 2), "S_fwd": S_hw,
            "storage": "f32",
        }

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(vol, P, sp, og, cfg, args)

    del vol
    torch.cuda.empty_cache()
    if not args.no_serve:
        out["fdk_serving"] = serve_bench(args)
    if not args.no_batched:
        torch.cuda.empty_cache()
        out["uneven_batch"] = batched_bench(args)
    if not args.no_configs:
        torch.cuda.empty_cache()
        out["configs"] = configs_bench(args)

    print(json.dumps(out), flush=True)
    return 0


def serve_bench(args):
    """FDK serving: seeded Poisson arrivals of 40% fan 512^2 / 360 views, 40% cone
    256^3 / 360 views (400x400 detector), 20% cone 512^3 / 720 views (1024x768);
    each request carries a noisy sinogram in pinned host memory and goes through
    cosine weighting, Ram-Lak filtering and backprojection; continuous batching
    over `serve_slots` streams with HIP graphs, admission by an HBM budget.
    Reported for FP32 and BF16 storage."""
    import numpy as np
    import torch

    import pyronn_layers as pl
    from pyronn_layers import geometry as geo
    from pyronn_layers import phantoms as ph
    from pyronn_layers.fdk import ConeFDK, FanFBP
    from pyronn_layers.serving import GpuExecutor, Request, Scheduler, latency_report, poisson_arrivals

    pl.load()
    dev = torch.device("cuda:0")
    ops = torch.ops.pyronn_layers
    geoms = {
        "fan512": geo.Geometry2D(volume_shape=(512, 512), volume_spacing=(0.5, 0.5), detector_shape=725,
                                 detector_spacing=1.0, n_views=360, angular_range=2 * math.pi, sid=1000.0,
                                 sdd=1200.0),
        "cone256": geo.ConeGeometry((256, 256, 256), (1.0, 1.0, 1.0), (400, 400), (1.0, 1.0), 360, 750.0, 1200.0),
        "cone512": geo.ConeGeometry((512, 512, 512), (1.0, 1.0, 1.0), (768, 1024), (1.2, 1.2), 720, 750.0, 1200.0),
    }
    mix = {"fan512": 0.4, "cone256": 0.4, "cone512": 0.2}
    rng = np.random.default_rng(1234)
    clean = {}
    log("serving: payloads")
    for k, g in geoms.items():
        if k.startswith("fan"):
            e = ph.random_ellipses(11, 16, g.volume_shape, g.volume_spacing)
            img = torch.tensor(ph.rasterise_ellipses(e, g.volume_shape, g.volume_spacing, g.volume_origin)[None],
                               device=dev)
            rays = torch.tensor(geo.fan_central_ray_vectors(g.n_views)[None], dtype=torch.float32, device=dev)
            one = lambda v: torch.full((1, 1), v, dtype=torch.float32, device=dev)
            sino = ops.fan_projection2d(img, [g.n_views, g.detector_shape],
                                        torch.tensor(g.volume_origin[None], dtype=torch.float32, device=dev),
                                        one(g.detector_origin),
                                        torch.tensor(np.asarray(g.volume_spacing)[None], dtype=torch.float32,
                                                     device=dev), one(g.detector_spacing), one(g.sid), one(g.sdd),
                                        rays, 0.5)
        else:
            e = ph.random_ellipsoids(12, 20, g.volume_shape, g.volume_spacing,
                                     axes_range_vox=(8.0, 100.0 * g.volume_shape[0] / 256))
            vol = ph.rasterise_ellipsoids_torch(e, g.volume_shape, g.volume_spacing, g.volume_origin, device=dev)
            sino = ops.cone_projection3d(vol[None], list(g.projection_shape),
                                         torch.tensor(g.volume_origin[None], dtype=torch.float32, device=dev),
                                         torch.tensor(np.asarray(g.volume_spacing)[None], dtype=torch.float32,
                                                      device=dev),
                                         torch.tensor(g.projection_matrices()[None], dtype=torch.float32,
                                                      device=dev),
                                         torch.full((1, 1), g.step_size, device=dev), torch.ones((1, 1), device=dev),
                                         False)
            del vol
        clean[k] = sino
    out = {"workload": "Poisson arrivals (seed 5, %.1f req/s, %d requests): 40%% fan 512^2/360 views, 40%% cone "
                       "256^3/360 views 400x400, 20%% cone 512^3/720 views 1024x768; noisy sinograms "
                       "(sigma = 0.01 max) from pinned host memory; latency = arrival -> volume in host memory"
                       % (args.serve_rate, args.serve_requests),
           "slots": args.serve_slots, "hip_graphs": True,
           "cone512_view_ranges": args.serve_chunks if args.serve_chunks > 1 else None,
           "cone256_view_ranges": args.serve_chunks_256 if args.serve_chunks_256 > 1 else None}
    for name, storage in (("f32", torch.float32), ("bf16", torch.bfloat16)):
        payloads = {}
        for k, s_ in clean.items():
            payloads[k] = []
            for _ in range(2):
                noisy = s_ + torch.randn(s_.shape, device=dev, generator=torch.Generator(device=dev).manual_seed(
                    int(rng.integers(1 << 30)))) * (0.01 * float(s_.max()))
                payloads[k].append(noisy.to(storage).cpu().pin_memory())
        plans = {k: (FanFBP(g, storage=storage) if k.startswith("fan") else ConeFDK(g, storage=storage))
                 for k, g in geoms.items()}
        # the ramp filter alone on the whole 512^3 request (720 views x 768 rows x 1024), resident in HBM:
        # the fused FFT kernel against the Toeplitz-GEMM filter; bytes = sinogram in + filtered out
        p512 = plans["cone512"]
        x = clean["cone512"].to(storage)
        filt = {}
        for impl in ("fft", "gemm"):
            p512.filter_impl = impl
            filt[impl + "_ms"] = round(_time_events(lambda: p512.filter(x), 5) * 1e3, 3)
        p512.filter_impl = "fft"
        nbytes = 2 * x.numel() * x.element_size()
        filt["fft_GBps"] = round(nbytes / (filt["fft_ms"] * 1e-3) / 1e9, 1)
        filt["fft_hbm_frac"] = round(filt["fft_GBps"] / 8000.0, 4)
        log(f"serving {name}: 512^3 ramp filter FFT {filt['fft_ms']:.2f} ms, GEMM {filt['gemm_ms']:.2f} ms")
        p512.ramp = None
        del x
        chunked = _chunked(args)
        ex = GpuExecutor(plans, slots=args.serve_slots, use_graphs=True, chunked=chunked,
                         group_kinds=("cone256",))
        log(f"serving {name}: capturing graphs")
        ex.warm()

        def serve(rate, group_max=1, at_once=False):
            reqs = poisson_arrivals(5, rate, args.serve_requests, mix)
            for r in reqs:
                r.payload = payloads[r.kind][r.rid % 2]
                if at_once:
                    r.arrival = 0.0
            res = Scheduler(ex, slots=args.serve_slots, hbm_budget=160 << 30, group_max=group_max).run(
                reqs, idle=lambda dt: time.sleep(dt) if dt > 0 else None)
            return _rounded(latency_report(res))

        log(f"serving {name}: {args.serve_requests} requests at {args.serve_rate} req/s")
        rep = serve(args.serve_rate)
        log(f"serving {name}: {rep['req_per_s']:.2f} req/s p50 {rep['p50_ms']:.1f} ms p99 {rep['p99_ms']:.1f} ms")
        out[name] = rep
        # p50 / p99 against the offered load, up to saturation
        curve = []
        for rate in args.serve_curve:
            r_ = serve(rate)
            curve.append({"offered_req_per_s": rate, **{k: r_[k] for k in ("req_per_s", "p50_ms", "p99_ms")},
                          "per_kind": r_["per_kind"]})
            log(f"serving {name} at {rate} req/s: {r_['req_per_s']:.2f} req/s p50 {r_['p50_ms']:.1f} ms "
                f"p99 {r_['p99_ms']:.1f} ms")
        out[name]["curve"] = curve
        # capacity: the same requests, all queued at t = 0 (throughput limited by the GPU, not arrivals),
        # each request alone in its slot, then queued cone 256^3 requests grouped into one batched
        # backprojection launch (up to 4 per group)
        for key, gm in (("capacity", 1), ("capacity_grouped_cone256", 4)):
            r_ = serve(args.serve_rate, group_max=gm, at_once=True)
            log(f"serving {name} saturated, group_max {gm}: {r_['req_per_s']:.2f} req/s")
            out[name][key] = {"req_per_s": r_["req_per_s"], "span_s": r_["span_s"], "completed": r_["completed"],
                              "p99_ms": r_["p99_ms"], "group_max": gm}
        out[name]["filter_cone512"] = filt
        del ex, plans, payloads
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    if args.serve_projections:
        out["with_projection_requests"] = serve_mixed_bench(args, geoms, clean)
    return out


def _chunked(args):
    """Request classes reconstructed in view ranges (uploads overlapping the reconstruction)."""
    out = {}
    if args.serve_chunks > 1:
        out["cone512"] = args.serve_chunks
    if args.serve_chunks_256 > 1:
        out["cone256"] = args.serve_chunks_256
    return out


def _rounded(rep):
    out = {}
    for k, v in rep.items():
        if isinstance(v, float):
            out[k] = round(v, 3)
        elif isinstance(v, dict):
            out[k] = _rounded(v)
        else:
            out[k] = v
    return out


def serve_mixed_bench(args, geoms, clean):
    """Reconstruction and projection requests in one Poisson stream: 30% fan 512^2 FBP,
    30% cone 256^3 FDK, 10% cone 512^3 FDK (view ranges), 15% cone 256^3 forward projections
    (volume in, 360 x 400 x 400 sinogram out) and 15% fan 512^2 forward projections
    (360 x 725), every geometry resident on the device; FP32 and BF16 storage."""
    import numpy as np
    import torch

    from pyronn_layers import phantoms as ph
    from pyronn_layers.fdk import ConeFDK, FanFBP
    from pyronn_layers.projectors import ConeProjector, FanProjector
    from pyronn_layers.serving import GpuExecutor, Scheduler, latency_report, poisson_arrivals

    dev = torch.device("cuda:0")
    gc, gf = geoms["cone256"], geoms["fan512"]
    e = ph.random_ellipsoids(12, 20, gc.volume_shape, gc.volume_spacing, axes_range_vox=(8.0, 100.0))
    vol = ph.rasterise_ellipsoids_torch(e, gc.volume_shape, gc.volume_spacing, gc.volume_origin, device=dev)[None]
    ef = ph.random_ellipses(11, 16, gf.volume_shape, gf.volume_spacing)
    img = torch.tensor(ph.rasterise_ellipses(ef, gf.volume_shape, gf.volume_spacing, gf.volume_origin)[None],
                       device=dev)
    mix = {"fan512": 0.3, "cone256": 0.3, "cone512": 0.1, "cone256_proj": 0.15, "fan512_proj": 0.15}
    out = {"workload": "Poisson arrivals (seed 9, %.1f req/s, %d requests): 30%% fan 512^2 FBP, 30%% cone 256^3 "
                       "FDK, 10%% cone 512^3 FDK, 15%% cone 256^3 forward projections, 15%% fan 512^2 forward "
                       "projections; inputs from pinned host memory, outputs back to host"
                       % (args.serve_rate, args.serve_requests), "slots": args.serve_slots}
    rng = np.random.default_rng(77)
    for name, storage in (("f32", torch.float32), ("bf16", torch.bfloat16)):
        plans = {"fan512": FanFBP(gf, storage=storage), "cone256": ConeFDK(gc, storage=storage),
                 "cone512": ConeFDK(geoms["cone512"], storage=storage),
                 "cone256_proj": ConeProjector(gc, storage=storage), "fan512_proj": FanProjector(gf, storage=storage,
                                                                                                  step_factor=0.5)}
        payloads = {k: [(s_ + torch.randn(s_.shape, device=dev, generator=torch.Generator(device=dev).manual_seed(
            int(rng.integers(1 << 30)))) * (0.01 * float(s_.max()))).to(storage).cpu().pin_memory()]
                    for k, s_ in clean.items()}
        payloads["cone256_proj"] = [vol.to(storage).cpu().pin_memory()]
        payloads["fan512_proj"] = [img.to(storage).cpu().pin_memory()]
        ex = GpuExecutor(plans, slots=args.serve_slots, use_graphs=True, chunked=_chunked(args))
        ex.warm()
        reqs = poisson_arrivals(9, args.serve_rate, args.serve_requests, mix)
        for r in reqs:
            r.payload = payloads[r.kind][0]
        res = Scheduler(ex, slots=args.serve_slots, hbm_budget=160 << 30).run(
            reqs, idle=lambda dt: time.sleep(dt) if dt > 0 else None)
        rep = _rounded(latency_report(res))
        log(f"serving with projections {name}: {rep['req_per_s']:.2f} req/s p50 {rep['p50_ms']:.1f} ms "
            f"p99 {rep['p99_ms']:.1f} ms; " + ", ".join(f"{k} p50 {v['p50_ms']:.1f} p99 {v['p99_ms']:.1f}"
                                                         for k, v in rep["per_kind"].items()))
        out[name] = rep
        del ex, plans, payloads
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    return out


def filter_bench(args):
    """The FDK ramp filter alone on the sinogram of one 512^3 request (720 views x 768 x
    1024, seeded uniform values, resident in HBM): the fused FFT kernel and the
    Toeplitz-GEMM filter, FP32 and BF16 storage; bytes = sinogram in + filtered out."""
    import torch

    from pyronn_layers import geometry as geo
    from pyronn_layers.fdk import ConeFDK

    dev = torch.device("cuda:0")
    g = geo.ConeGeometry((512, 512, 512), (1.0, 1.0, 1.0), (768, 1024), (1.2, 1.2), 720, 750.0, 1200.0)
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    sino = torch.rand((1, g.n_views) + tuple(g.detector_shape), device=dev, generator=gen)
    out = {"workload": "cone 512^3 request: 720 views x 768 x 1024 detector rows"}
    for name, storage in (("f32", torch.float32), ("bf16", torch.bfloat16)):
        plan = ConeFDK(g, storage=storage)
        x = sino.to(storage)
        res = {}
        for impl in ("fft", "gemm"):
            plan.filter_impl = impl
            res[impl + "_ms"] = round(_time_events(lambda: plan.filter(x), 5) * 1e3, 3)
        nbytes = 2 * x.numel() * x.element_size()
        res["fft_GBps"] = round(nbytes / (res["fft_ms"] * 1e-3) / 1e9, 1)
        log(f"filter {name}: FFT {res['fft_ms']:.2f} ms, GEMM {res['gemm_ms']:.2f} ms")
        out[name] = res
        del plan, x
    return out


def _time_events(fn, reps):
    """Mean duration (s) of fn() over reps launches, HIP events on the launch stream."""
    import torch

    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / 1e3 / reps


def configs_bench(args):
    """The other benchmark configurations, each as forward + backprojection of its
    sinogram, FP32, with the same units as the headline line:
      C1 parallel 256^2 (1 mm), 180 views over pi, 363 detector pixels, step 0.2 (the
         reference's fixed factor), 12 seeded ellipses;
      C2 fan 512^2 (0.5 mm), 360 views over 2 pi, 725 pixels, SID 1000, SDD 1200, step 0.5;
      C3 cone 256^3 (1 mm), 360 views over 2 pi, 400 x 400 (1 mm), SID 750, SDD 1200,
         step 0.5, software and hardware_interp.
    Gsamples/s = (forward samples + backprojection updates) / (t_fwd + t_bp); the 2D ops
    are also timed on batches of 32 images (throughput regime).  roofline_frac uses the
    algorithmic bytes (2D: 4 neighbours per forward sample, 2 per backprojection update;
    3D: 8 and 4) over 8 TB/s.  The CPU oracle is timed on the same problems (C1, C2 in full;
    C3 on 8 evenly spaced views and 8 central slices)."""
    import numpy as np
    import torch

    import oracle as orc
    import pyronn_layers as pl
    from pyronn_layers import geometry as geo
    from pyronn_layers import metrics
    from pyronn_layers import phantoms as ph

    pl.load()
    dev = torch.device("cuda:0")
    ops = torch.ops.pyronn_layers
    res = {}
    t1 = lambda v: torch.full((1, 1), float(v), device=dev)

    def twod(name, fan, H, W, s, n, arc, det, ds, sid, sdd, step, n_ell, seed):
        org = geo.centred_origin((H, W), s)
        e = ph.random_ellipses(seed, n_ell, (H, W), s)
        img = ph.rasterise_ellipses(e, (H, W), s, org).astype(np.float32)
        rays = (geo.fan_central_ray_vectors(n, arc) if fan else geo.parallel_ray_vectors(n, arc)).astype(np.float32)
        do = -(det - 1) / 2 * ds
        fwd = ops.fan_projection2d if fan else ops.parallel_projection2d
        bwd = ops.fan_backprojection2d if fan else ops.parallel_backprojection2d
        out = {}
        for B in (1, 32):
            im = torch.tensor(img, device=dev)[None].repeat(B, 1, 1).contiguous()
            g = [torch.tensor(org[None], dtype=torch.float32, device=dev).repeat(B, 1), t1(do).repeat(B, 1),
                 torch.tensor(np.asarray(s, np.float32)[None], device=dev).repeat(B, 1), t1(ds).repeat(B, 1)]
            if fan:
                g += [t1(sid).repeat(B, 1), t1(sdd).repeat(B, 1)]
            r = torch.tensor(rays[None], device=dev).repeat(B, 1, 1).contiguous()
            sino = fwd(im, [n, det], *g, r, step)
            tf = _time_events(lambda: fwd(im, [n, det], *g, r, step), 20)
            tb = _time_events(lambda: bwd(sino, [H, W], *g, r), 20)
            out[B] = (tf, tb)
        t0 = time.perf_counter()
        f = orc.fan_forward if fan else orc.parallel_forward
        b = orc.fan_backward if fan else orc.parallel_backward
        kw = dict(sid=sid, sdd=sdd) if fan else {}
        sref, _, S = f(img, rays, det, org, s, do, ds, step_factor=step, **kw)
        cf = time.perf_counter() - t0
        t0 = time.perf_counter()
        b(sref, rays, (H, W), org, s, do, ds, **kw)
        cb = time.perf_counter() - t0
        U = H * W * n
        Bytes = S * 4 * 4 + n * det * 4 + U * 2 * 4 + H * W * 4
        tf, tb = out[1]
        tf32, tb32 = out[32]
        res[name] = {
            "t_fwd_ms": round(tf * 1e3, 4), "t_bp_ms": round(tb * 1e3, 4), "S_fwd": int(S), "U_bp": U,
            "gsamples_s": round((S + U) / (tf + tb) / 1e9, 3),
            "gsamples_s_batch32": round(32 * (S + U) / (tf32 + tb32) / 1e9, 3),
            "roofline_frac_batch32": round(32 * Bytes / (tf32 + tb32) / HBM_PEAK, 4),
            "cpu_oracle_gsamples_s": round((S + U) / (cf + cb) / 1e9, 4), "cpu_threads": orc.num_threads(),
        }
        log(f"configs {name}: GPU {res[name]['gsamples_s']} Gsamples/s (batch 32: {res[name]['gsamples_s_batch32']}), "
            f"CPU {res[name]['cpu_oracle_gsamples_s']}")

    log("configs: C1 parallel")
    twod("C1_parallel_256", False, 256, 256, (1.0, 1.0), 180, math.pi, 363, 1.0, None, None, 0.2, 12, 1)
    log("configs: C2 fan")
    twod("C2_fan_512", True, 512, 512, (0.5, 0.5), 360, 2 * math.pi, 725, 1.0, 1000.0, 1200.0, 0.5, 16, 2)

    log("configs: C3 cone")
    shape, det, n = (256, 256, 256), (400, 400), 360
    sp = np.ones(3, np.float32)
    og = geo.centred_origin(shape, sp).astype(np.float32)
    P = geo.circular_projection_matrices(n, 750.0, 1200.0, det, (1.0, 1.0)).astype(np.float32)
    e = ph.random_ellipsoids(3, 20, shape, sp, axes_range_vox=(8.0, 100.0))
    vol = ph.rasterise_ellipsoids_torch(e, shape, sp, og, device=dev)[None].contiguous()
    geo_t = (torch.tensor(og[None], device=dev), torch.tensor(sp[None], device=dev), torch.tensor(P[None], device=dev),
             t1(0.5), t1(1.0))
    U = int(np.prod(shape)) * n
    for hw in (False, True):
        S = metrics.cone_forward_sample_count(P, shape, sp, og, det, 0.5, hardware_interp=hw, device=dev)
        sino = ops.cone_projection3d(vol, [n, *det], *geo_t, hw)
        tf = _time_events(lambda: ops.cone_projection3d(vol, [n, *det], *geo_t, hw), 5)
        tb = _time_events(lambda: ops.cone_backprojection3d(sino, list(shape), *geo_t, hw), 5)
        Bytes = S * 8 * 4 + n * det[0] * det[1] * 4 + U * 4 * 4 + int(np.prod(shape)) * 4
        res["C3_cone_256" + ("_hw" if hw else "")] = {
            "t_fwd_ms": round(tf * 1e3, 3), "t_bp_ms": round(tb * 1e3, 3), "S_fwd": int(S), "U_bp": U,
            "gsamples_s": round((S + U) / (tf + tb) / 1e9, 3),
            "roofline_frac": round(Bytes / (tf + tb) / HBM_PEAK, 4),
        }
        log(f"configs C3 hw={hw}: fwd {tf * 1e3:.2f} ms bp {tb * 1e3:.2f} ms")
    # CPU oracle on a bounded sample of C3 (software interpolation)
    vh = vol[0].cpu().numpy()
    inv, src, _, _, _ = orc.cone_setup(P, sp, og)
    views = np.linspace(0, n, 8, endpoint=False).astype(int)
    t0 = time.perf_counter()
    Sc = 0
    for v in views:
        _, _, ns = orc.cone_forward(vh, inv[v:v + 1], src[v:v + 1], det, sp, 0.5)
        Sc += ns
    cf = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.cone_backward(ops.cone_projection3d(vol, [n, *det], *geo_t, False)[0].cpu().numpy(), P, shape, sp, og,
                      zrange=(124, 132))
    cb = time.perf_counter() - t0
    res["C3_cone_256"]["cpu_oracle_gsamples_s"] = round((Sc + 8 * 256 * 256 * n) / (cf + cb) / 1e9, 4)
    res["C3_cone_256"]["cpu_sample"] = "8 of 360 forward views + 8 of 256 backprojection slices"
    return res


def batched_bench(args):
    """Eight uneven cone problems (128^3 ... 512^3, 360-720 views, detectors 256^2 ...
    1024x768, SID 600-1000 mm, each with its own seeded 20-ellipsoid phantom): forward
    and backprojection each in ONE launch with autotuned launch parameters,
    FP32 and BF16 storage; versus the same problems launched one after another; error
    of the GPU forward against the FP64 oracle on views of the smallest problem."""
    import numpy as np
    import torch

    import oracle as orc
    import pyronn_layers as pl
    from pyronn_layers import batched, metrics
    from pyronn_layers import geometry as geo
    from pyronn_layers import phantoms as ph
    from pyronn_layers.autotune import Autotuner

    pl.load()
    dev = torch.device("cuda:0")
    sizes = [128, 192, 256, 256, 320, 384, 448, 512]
    views = [360, 400, 450, 500, 560, 620, 680, 720]
    sids = np.linspace(600.0, 1000.0, len(sizes))
    geoms, vols = [], []
    log("uneven batch: phantoms")
    for k, (N, n, sid) in enumerate(zip(sizes, views, sids)):
        det = (max(256, N * 3 // 2), 2 * N)
        g = geo.ConeGeometry((N, N, N), (1.0, 1.0, 1.0), det, (1.2, 1.2), n, float(sid), float(sid) * 1.6)
        e = ph.random_ellipsoids(100 + k, 20, g.volume_shape, g.volume_spacing,
                                 axes_range_vox=(8.0 * N / 256, 100.0 * N / 256))
        geoms.append(g)
        vols.append(ph.rasterise_ellipsoids_torch(e, g.volume_shape, g.volume_spacing, g.volume_origin, device=dev))
    S = sum(metrics.cone_forward_sample_count(g.projection_matrices(), g.volume_shape, g.volume_spacing,
                                              g.volume_origin, g.detector_shape, g.step_size, device=dev)
            for g in geoms)
    U = sum(metrics.cone_backprojection_update_count(g.volume_shape, g.n_views) for g in geoms)
    tuner = Autotuner(cache_path=None, reps=1)
    # geometry arrays resident on the device, as a serving caller keeps them: the timed calls
    # upload no projection matrices
    dgeo = batched.device_geometry(geoms, dev)

    def timed(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        r = fn()
        b.record()
        torch.cuda.synchronize()
        return r, a.elapsed_time(b) / 1e3

    out = {"workload": "8 cone problems: " + ", ".join(f"{g.volume_shape[0]}^3/{g.n_views}v/{g.detector_shape[1]}x"
                                                      f"{g.detector_shape[0]}" for g in geoms),
           "S_fwd": S, "U_bp": U}
    for name, storage in (("f32", torch.float32), ("bf16", torch.bfloat16)):
        vin = [v.to(storage) for v in vols]
        log(f"uneven batch {name}: tuning")
        tf = tuner.forward(vin, geoms)
        sinos, t_fwd = timed(lambda: batched.cone_projection3d_batched(vin, geoms, blocks_per_cu=tf["blocks_per_cu"],
                                                                       forward_tile=tf["forward_tile"],
                                                                       device_geometry=dgeo))
        sin = [s_.to(storage) for s_ in sinos]
        tb = tuner.backward(sin, geoms)
        _, t_bp = timed(lambda: batched.cone_backprojection3d_batched(sin, geoms, z_per_lane=tb["z_per_lane"],
                                                                      blocks_per_cu=tb["blocks_per_cu"],
                                                                      device_geometry=dgeo))
        # the same problems, one batched launch per problem, one after another
        def seq_fwd():
            for v, g, d in zip(vin, geoms, dgeo):
                batched.cone_projection3d_batched([v], [g], blocks_per_cu=tf["blocks_per_cu"],
                                                  forward_tile=tf["forward_tile"], device_geometry=[d])
        def seq_bp():
            for s_, g, d in zip(sin, geoms, dgeo):
                batched.cone_backprojection3d_batched([s_], [g], z_per_lane=tb["z_per_lane"], device_geometry=[d])
        _, ts_f = timed(seq_fwd)
        _, ts_b = timed(seq_bp)
        # error against the FP64 oracle: two views of the 128^3 problem, forward
        g0 = geoms[0]
        inv, src, inv64, src64, _ = orc.cone_setup(g0.projection_matrices().astype(np.float32),
                                                   np.float32(g0.volume_spacing), g0.volume_origin.astype(np.float32))
        r64, _ = orc.cone_forward_f64(vols[0].double().cpu().numpy(), inv64, src64, g0.detector_shape,
                                      np.asarray(g0.volume_spacing), g0.step_size, views=(0, 2))
        got = sinos[0][:2].double().cpu().numpy()
        err = float(np.linalg.norm(got - r64) / np.linalg.norm(r64))
        log(f"uneven batch {name}: fwd {t_fwd * 1e3:.1f} ms bp {t_bp * 1e3:.1f} ms (sequential {ts_f * 1e3:.1f} / "
            f"{ts_b * 1e3:.1f} ms), rel L2 vs FP64 {err:.2e}")
        out[name] = {"gsamples_s": round((S + U) / (t_fwd + t_bp) / 1e9, 2), "t_fwd_ms": round(t_fwd * 1e3, 2),
                     "t_bp_ms": round(t_bp * 1e3, 2), "sequential_t_fwd_ms": round(ts_f * 1e3, 2),
                     "sequential_t_bp_ms": round(ts_b * 1e3, 2),
                     "tuned_fwd": {"blocks_per_cu": tf["blocks_per_cu"], "tile": tf["tile_shape"],
                                   "trials_ms": {k: round(v, 2) for k, v in tf["trials_ms"].items()}},
                     "tuned_bp": [tb["z_per_lane"], tb["blocks_per_cu"]],
                     "rel_l2_vs_fp64_fwd_views_0_1_of_128^3": err}
        del sinos, sin, vin
        torch.cuda.empty_cache()
    return out


def precision_vs_fp64(vol_t, sino16_t, rec16_t, storage, P, sp, og, cfg, views=(0, 180), slices=(128, 256)):
    """Relative L2 errors of a 16-bit-storage run against the FP64 oracle, on two forward
    views and two backprojected z slices of the 512^3 workload:
    * kernel: FP64 fed the same dequantised input (the 16-bit volume; the 16-bit sinogram
      the backprojection read);
    * end_to_end: FP64 fed the input before rounding to 16 bits (the FP32 phantom; the
      FP32 sinogram that the 16-bit forward produced)."""
    import numpy as np
    import torch

    import oracle as orc

    def rel(a, b):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    D, H, W = cfg["shape"]
    t0 = time.perf_counter()
    _, _, inv64, src64, _ = orc.cone_setup(P, sp, og)
    v32 = vol_t[0].float().cpu().numpy()
    vq = vol_t[0].to(storage).float().cpu().numpy()
    got = sino16_t[0].cpu().numpy()
    k_f, e_f = [], []
    for v in views:
        ref_q, _ = orc.cone_forward_f64(vq, inv64, src64, cfg["det"], sp, cfg["step"], views=(v, v + 1))
        ref_e, _ = orc.cone_forward_f64(v32, inv64, src64, cfg["det"], sp, cfg["step"], views=(v, v + 1))
        k_f.append(rel(got[v:v + 1], ref_q))
        e_f.append(rel(got[v:v + 1], ref_e))
    # the backprojection read sino16.to(storage): FP64 on that (kernel) and on the FP32 sinogram
    s32 = sino16_t[0].cpu().numpy()
    sq = sino16_t[0].to(storage).float().cpu().numpy()
    rec = rec16_t[0].cpu().numpy()
    k_b, e_b = [], []
    for z in slices:
        ref_q = orc.cone_backward_f64(sq, P, (D, H, W), sp, og, zrange=(z, z + 1))
        ref_e = orc.cone_backward_f64(s32, P, (D, H, W), sp, og, zrange=(z, z + 1))
        k_b.append(rel(rec[z:z + 1], ref_q))
        e_b.append(rel(rec[z:z + 1], ref_e))
    log(f"{storage} error vs fp64: forward kernel {max(k_f):.2e} end-to-end {max(e_f):.2e}; "
        f"backprojection kernel {max(k_b):.2e} end-to-end {max(e_b):.2e} ({time.perf_counter() - t0:.1f} s)")
    return {
        "forward_views": list(views), "backprojection_slices": list(slices),
        "forward_kernel_rel_l2": max(k_f), "forward_end_to_end_rel_l2": max(e_f),
        "backprojection_kernel_rel_l2": max(k_b), "backprojection_end_to_end_rel_l2": max(e_b),
        "bar": 1e-3,
    }


def host_cpu():
    """Host description for the CPU baseline record."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"nproc": affinity, "os_cpu_count": os.cpu_count(), "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
            "cpu_model": model}


def cpu_baseline(vol_t, P, sp, og, cfg, args):
    """The f32 oracle (the reference kernel arithmetic on the host, OpenMP) on a
    bounded sample: `cpu_views` evenly spaced forward views and the backprojection
    of `cpu_slices` central z slices over all views."""
    import numpy as np
    import torch

    import oracle as orc
    import pyronn_layers as pl

    D, H, W = cfg["shape"]
    n = cfg["n"]
    Hd, Wd = cfg["det"]
    vol = vol_t[0].float().cpu().numpy()
    inv, src, _, _, _ = orc.cone_setup(P, sp, og)
    views = np.linspace(0, n, args.cpu_views, endpoint=False).astype(int)
    log(f"cpu baseline: {len(views)} forward views, {args.cpu_slices} bp slices, {orc.num_threads()} threads")
    t0 = time.perf_counter()
    S = 0
    sub_sino = []
    for v in views:
        o, _, ns = orc.cone_forward(vol, inv[v:v + 1], src[v:v + 1], (Hd, Wd), sp, cfg["step"])
        S += ns
        sub_sino.append(o)
    tf = time.perf_counter() - t0
    # Backprojection sample: full sinogram from the GPU (not timed), central slab.
    ops = torch.ops.pyronn_layers
    dev = vol_t.device
    sino = ops.cone_projection3d(vol_t, [n, Hd, Wd], torch.tensor(og[None], device=dev),
                                 torch.tensor(sp[None], device=dev), torch.tensor(P[None], device=dev),
                                 torch.full((1, 1), cfg["step"], device=dev), torch.ones((1, 1), device=dev),
                                 False)[0].cpu().numpy()
    z0 = D // 2 - args.cpu_slices // 2
    t0 = time.perf_counter()
    orc.cone_backward(sino, P, (D, H, W), sp, og, zrange=(z0, z0 + args.cpu_slices))
    tb = time.perf_counter() - t0
    U = args.cpu_slices * H * W * n
    return {
        "value": round((S + U) / (tf + tb) / 1e9, 4), "unit": "Gsamples/s", "cores": orc.num_threads(),
        "kind": "port", **host_cpu(),
        "sample": f"f32 oracle: forward of {len(views)} of {n} views ({S} samples, {tf:.1f} s) + backprojection "
                  f"of {args.cpu_slices} of {D} z-slices over all views ({U} updates, {tb:.1f} s)",
    }


if __name__ == "__main__":
    sys.exit(main())
