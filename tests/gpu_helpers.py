"""Numpy-in / numpy-out wrappers around the C ABI (ctypes) for the GPU tests.

Device memory and the stream come from torch; every compute call goes through
libpyronn_mi.so's extern "C" entry points.
"""
import ctypes

import numpy as np
import torch

from pyronn_layers import native as nat

DT = {"f32": (nat.F32, torch.float32), "bf16": (nat.BF16, torch.bfloat16), "f16": (nat.F16, torch.float16)}


def dev(a, dtype=torch.float32):
    return torch.from_numpy(np.array(a, copy=True, order="C")).to("cuda").to(dtype).contiguous()


def p(t):
    return ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _batched(a, per_item_ndim):
    a = np.asarray(a)
    return a if a.ndim > per_item_ndim else a[None]


def quantise(a, storage):
    """Values the kernels see for `storage` (dequantised back to float32)."""
    return dev(a, DT[storage][1]).float().cpu().numpy()


def cone_setup(P, spacing_zyx, origin_zyx):
    P = _batched(np.asarray(P, np.float32).reshape(-1, 3, 4) if np.asarray(P).ndim <= 3 else P, 3)
    B, n = P.shape[0], P.shape[1]
    Pd = dev(P)
    sp = dev(_batched(spacing_zyx, 1))
    og = dev(_batched(origin_zyx, 1))
    inv = torch.empty((B, n, 9), device="cuda")
    src = torch.empty((B, n, 3), device="cuda")
    nat.call("pyronn_cone_geometry_setup", p(Pd), p(og), p(sp), B, n, p(inv), p(src), stream())
    torch.cuda.synchronize()
    return inv.cpu().numpy(), src.cpu().numpy()


def cone_forward(vol, P, spacing_zyx, origin_zyx, det_shape, step, hw=False, storage="f32", mult=1.0):
    vol = _batched(vol, 3)
    B, D, H, W = vol.shape
    P = np.asarray(P, np.float32)
    P = P if P.ndim == 4 else P[None].repeat(B, 0)
    n = P.shape[1]
    Hd, Wd = det_shape
    code, tdt = DT[storage]
    v = dev(vol, tdt)
    Pd = dev(P)
    sp = dev(np.broadcast_to(np.asarray(spacing_zyx, np.float32).reshape(-1, 3), (B, 3)))
    og = dev(np.broadcast_to(np.asarray(origin_zyx, np.float32).reshape(-1, 3), (B, 3)))
    st = dev(np.broadcast_to(np.asarray(step, np.float32).reshape(-1, 1), (B, 1)))
    mu = dev(np.full((B, 1), mult, np.float32))
    out = torch.empty((B, n, Hd, Wd), device="cuda")
    ws = nat.lib().pyronn_cone_projection3d_workspace_size(B, D, H, W, n, code)
    work = torch.empty(ws, dtype=torch.uint8, device="cuda")
    nat.call("pyronn_cone_projection3d", p(v), code, B, D, H, W, p(og), p(sp), p(Pd), p(st), p(mu), n, Hd, Wd,
             int(bool(hw)), p(out), p(work), ws, stream())
    torch.cuda.synchronize()
    return out.cpu().numpy()


def cone_backward(sino, P, spacing_zyx, origin_zyx, vol_shape, mult=1.0, hw=False, storage="f32", step=0.5):
    sino = _batched(sino, 3)
    B, n, Hd, Wd = sino.shape
    D, H, W = vol_shape
    P = np.asarray(P, np.float32)
    P = P if P.ndim == 4 else P[None].repeat(B, 0)
    code, tdt = DT[storage]
    s_ = dev(sino, tdt)
    Pd = dev(P)
    sp = dev(np.broadcast_to(np.asarray(spacing_zyx, np.float32).reshape(-1, 3), (B, 3)))
    og = dev(np.broadcast_to(np.asarray(origin_zyx, np.float32).reshape(-1, 3), (B, 3)))
    st = dev(np.full((B, 1), step, np.float32))
    mu = dev(np.broadcast_to(np.asarray(mult, np.float32).reshape(-1, 1), (B, 1)))
    out = torch.empty((B, D, H, W), device="cuda")
    ws = nat.lib().pyronn_cone_backprojection3d_workspace_size(B, n, Hd, Wd, code)
    work = torch.empty(ws, dtype=torch.uint8, device="cuda")
    nat.call("pyronn_cone_backprojection3d", p(s_), code, B, n, Hd, Wd, p(og), p(sp), p(Pd), p(st), p(mu), D, H,
             W, int(bool(hw)), p(out), p(work), ws, stream())
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _geo2d(B, vo, do_, vs, ds, rays, n, sid=None, sdd=None):
    g = dict(
        vo=dev(np.broadcast_to(np.asarray(vo, np.float32).reshape(-1, 2), (B, 2))),
        do=dev(np.broadcast_to(np.asarray(do_, np.float32).reshape(-1, 1), (B, 1))),
        vs=dev(np.broadcast_to(np.asarray(vs, np.float32).reshape(-1, 2), (B, 2))),
        ds=dev(np.broadcast_to(np.asarray(ds, np.float32).reshape(-1, 1), (B, 1))),
        rays=dev(np.broadcast_to(np.asarray(rays, np.float32).reshape(-1, n, 2), (B, n, 2))),
    )
    if sid is not None:
        g["sid"] = dev(np.broadcast_to(np.asarray(sid, np.float32).reshape(-1, 1), (B, 1)))
        g["sdd"] = dev(np.broadcast_to(np.asarray(sdd, np.float32).reshape(-1, 1), (B, 1)))
    return g


def forward2d(fan, vol, rays, det, vo, vs, do_, ds, sid=None, sdd=None, step_factor=0.2, storage="f32"):
    vol = _batched(vol, 2)
    B, H, W = vol.shape
    n = np.asarray(rays).reshape(-1, 2).shape[0] if np.asarray(rays).ndim <= 2 else np.asarray(rays).shape[1]
    code, tdt = DT[storage]
    v = dev(vol, tdt)
    g = _geo2d(B, vo, do_, vs, ds, rays, n, sid, sdd)
    out = torch.empty((B, n, det), device="cuda")
    if fan:
        nat.call("pyronn_fan_projection2d", p(v), code, B, H, W, p(g["vo"]), p(g["do"]), p(g["vs"]), p(g["ds"]),
                 p(g["sid"]), p(g["sdd"]), p(g["rays"]), n, det, float(step_factor), p(out), stream())
    else:
        nat.call("pyronn_parallel_projection2d", p(v), code, B, H, W, p(g["vo"]), p(g["do"]), p(g["vs"]),
                 p(g["ds"]), p(g["rays"]), n, det, float(step_factor), p(out), stream())
    torch.cuda.synchronize()
    return out.cpu().numpy()


def backward2d(fan, sino, rays, vol_shape, vo, vs, do_, ds, sid=None, sdd=None, storage="f32"):
    sino = _batched(sino, 2)
    B, n, det = sino.shape
    H, W = vol_shape
    code, tdt = DT[storage]
    s_ = dev(sino, tdt)
    g = _geo2d(B, vo, do_, vs, ds, rays, n, sid, sdd)
    out = torch.empty((B, H, W), device="cuda")
    if fan:
        nat.call("pyronn_fan_backprojection2d", p(s_), code, B, n, det, p(g["vo"]), p(g["do"]), p(g["vs"]),
                 p(g["ds"]), p(g["sid"]), p(g["sdd"]), p(g["rays"]), H, W, p(out), stream())
    else:
        nat.call("pyronn_parallel_backprojection2d", p(s_), code, B, n, det, p(g["vo"]), p(g["do"]), p(g["vs"]),
                 p(g["ds"]), p(g["rays"]), H, W, p(out), stream())
    torch.cuda.synchronize()
    return out.cpu().numpy()


def parity(gpu, ref, rel_l2=1e-5, max_rel=1e-4):
    """The FP32 bar: relative L2 and max |diff| relative to max |ref|."""
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    assert gpu.shape == ref.shape, (gpu.shape, ref.shape)
    assert np.all(np.isfinite(gpu)), "non-finite output"
    nrm = np.linalg.norm(ref)
    l2 = np.linalg.norm(gpu - ref) / (nrm if nrm > 0 else 1.0)
    mx = np.abs(gpu - ref).max() / max(np.abs(ref).max(), 1e-30)
    exact = float(np.mean(gpu == ref))
    assert l2 <= rel_l2 and mx <= max_rel, f"rel L2 {l2:.3e} (<= {rel_l2}), max {mx:.3e} (<= {max_rel}), exact {exact:.4f}"
    return l2, mx, exact
