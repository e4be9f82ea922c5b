"""Throughput of the cone-beam projector / backprojector pair on one MI355X.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]

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
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-bf16", action="store_true", help="skip the BF16-storage run")
    ap.add_argument("--cpu-views", type=int, default=4, help="forward views in the CPU sample")
    ap.add_argument("--cpu-slices", type=int, default=16, help="z slices in the CPU backprojection sample")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one plain 1 GiB device copy (reference for HBM byte counters)")
    ap.add_argument("--bf16-only", action="store_true", help="time only the BF16-storage run")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path.",
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

    def run(storage):
        v = vol.to(storage)
        times_f, times_b = [], []
        sino = None
        for it in range(args.warmup + args.steps):
            timed = it >= args.warmup
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            torch.cuda.synchronize()
            e0.record()
            sino = ops.cone_projection3d(v, [n, Hd, Wd], vo, vs, Pt, st, mu, False)
            e1.record()
            s_in = sino if storage == torch.float32 else sino.to(storage)
            e1b = torch.cuda.Event(enable_timing=True)
            e1b.record()
            rec = ops.cone_backprojection3d(s_in, [D, H, W], vo, vs, Pt, st, mu, False)
            e2.record()
            torch.cuda.synchronize()
            tf, tb = e0.elapsed_time(e1) / 1e3, e1b.elapsed_time(e2) / 1e3
            log(f"{'step' if timed else 'warmup'} {it} {str(storage)}: fwd {tf * 1e3:.1f} ms, bp {tb * 1e3:.1f} ms")
            if timed:
                times_f.append(tf)
                times_b.append(tb)
            del rec
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
            "achieved": round(B_dom / t_dom / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(B_dom / t_dom / HBM_PEAK, 4), "traffic": None,
        },
        "fwd_back": {
            "t_fwd_ms": round(tf * 1e3, 2), "t_bp_ms": round(tb * 1e3, 2),
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
        tf16, tb16, sino16 = run(torch.bfloat16)
        del sino16
        Bf = S_fwd * 8 * 2 + R * 4
        Bb = U_bp * 4 * 2 + V * 4
        g16 = (S_fwd + U_bp) / (tf16 + tb16) / 1e9
        out["bf16"] = {
            "value": round(g16, 3), "unit": "Gsamples/s", "speedup_vs_f32": round(g16 / gs, 3),
            "t_fwd_ms": round(tf16 * 1e3, 2), "t_bp_ms": round(tb16 * 1e3, 2),
            "hbm_roofline_frac_fwd_back": round((Bf + Bb) / (tf16 + tb16) / HBM_PEAK, 4),
            "storage": "bf16 volume and sinogram, fp32 accumulation, fp32 outputs",
        }
    else:
        del sino

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(vol, P, sp, og, cfg, args)

    print(json.dumps(out), flush=True)
    return 0


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
        "kind": "port",
        "sample": f"f32 oracle: forward of {len(views)} of {n} views ({S} samples, {tf:.1f} s) + backprojection "
                  f"of {args.cpu_slices} of {D} z-slices over all views ({U} updates, {tb:.1f} s)",
    }


if __name__ == "__main__":
    sys.exit(main())
