"""Uneven batches: cone projections / backprojections of problems with different
volume sizes, detectors and geometries in one launch each
(pyronn_cone_projection3d_batched, pyronn_cone_backprojection3d_batched: one workgroup per
item, or a persistent grid with blocks_per_cu > 0).

Every output equals the single-problem operator's output bit for bit; the
batched launch only changes which workgroup computes which tile.
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import numpy as np
import torch

from . import geometry as geo
from . import native as nat

_DT = {torch.float32: nat.F32, torch.bfloat16: nat.BF16, torch.float16: nat.F16}


def _f32(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32).to(dev).contiguous()


class _Geo:
    """Device copies of one ConeGeometry's arrays (P, vo, vs, step, mult; g the geometry)."""

    def __init__(self, g: geo.ConeGeometry, dev):
        self.g = g
        self.P = _f32(g.projection_matrices(), dev)
        self.vo = _f32(np.asarray(g.volume_origin), dev)
        self.vs = _f32(np.asarray(g.volume_spacing), dev)
        self.step = _f32([g.step_size], dev)
        self.mult = _f32([g.projection_multiplier], dev)


def _validate(inputs: Sequence[torch.Tensor], geometries: Sequence[geo.ConeGeometry], forward: bool):
    """The kernels index each input with the shapes of its geometry: refuse any tensor that
    is not a CUDA tensor on the first input's device with exactly those shapes."""
    if len(inputs) != len(geometries):
        raise ValueError(f"{len(inputs)} inputs for {len(geometries)} geometries")
    _check(len(inputs))
    dev = inputs[0].device
    for k, (x, g) in enumerate(zip(inputs, geometries)):
        if not isinstance(x, torch.Tensor) or not x.is_cuda:
            raise ValueError(f"input {k} is not a CUDA tensor")
        if x.device != dev:
            raise ValueError(f"input {k} is on {x.device}, input 0 on {dev}")
        want = tuple(g.volume_shape) if forward else (g.n_views,) + tuple(g.detector_shape)
        if tuple(x.shape) != want:
            what = "volume_shape" if forward else "(n_views,) + detector_shape"
            raise ValueError(f"input {k} has shape {tuple(x.shape)}, its geometry's {what} is {want}")
        if x.dtype not in _DT:
            raise TypeError(f"input {k}: storage dtype {x.dtype} not supported")


def _problems(inputs, outputs, geos: List[_Geo]):
    arr = (nat.ConeProblem * len(inputs))()
    for k, (x, y, gg) in enumerate(zip(inputs, outputs, geos)):
        g = gg.g
        D, H, W = g.volume_shape
        Hd, Wd = g.detector_shape
        arr[k] = nat.ConeProblem(x.data_ptr(), y.data_ptr(), gg.vo.data_ptr(), gg.vs.data_ptr(), gg.P.data_ptr(),
                                 gg.step.data_ptr(), gg.mult.data_ptr(), D, H, W, g.n_views, Hd, Wd)
    return arr


def _check(n):
    if not 1 <= n <= nat.MAX_BATCHED_PROBLEMS:
        raise ValueError(f"1..{nat.MAX_BATCHED_PROBLEMS} problems per batched launch, got {n}")


def device_geometry(geometries: Sequence[geo.ConeGeometry], device) -> list:
    """The device arrays of each geometry, to keep resident across batched calls
    (device_geometry= of cone_projection3d_batched / cone_backprojection3d_batched), so that a
    call uploads no projection matrices."""
    dev = torch.device(device)
    return [_Geo(g, dev) for g in geometries]


def cone_projection3d_batched(volumes: Sequence[torch.Tensor], geometries: Sequence[geo.ConeGeometry],
                              hardware_interp: bool = False, blocks_per_cu: int = 0,
                              forward_tile: int = 0, device_geometry=None) -> List[torch.Tensor]:
    """volumes[k] [D_k, H_k, W_k] (one storage dtype for all) -> sinograms [n_k, H_d,k, W_d,k] float32.
    blocks_per_cu and forward_tile (0..3: 16x16, 8x32, 32x8 and 16x16 with 8x8 waves) are launch
    shapes that never change a result.  device_geometry: optional per-problem device arrays
    already resident (as for cone_backprojection3d_batched)."""
    _validate(volumes, geometries, True)
    dev = volumes[0].device
    dt = volumes[0].dtype
    if any(v.dtype != dt for v in volumes):
        raise TypeError("all volumes of a batched launch share one storage dtype")
    vols = [v.contiguous() for v in volumes]
    geos = (_device_geometry(device_geometry, geometries, dev) if device_geometry is not None
            else [_Geo(g, dev) for g in geometries])
    outs = [torch.empty((g.n_views,) + tuple(g.detector_shape), dtype=torch.float32, device=dev) for g in geometries]
    arr = _problems(vols, outs, geos)
    L = nat.lib()
    ws_bytes = L.pyronn_cone_projection3d_batched_workspace_size(arr, len(vols), _DT[dt])
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    nat.call("pyronn_cone_projection3d_batched", arr, len(vols), _DT[dt], int(bool(hardware_interp)),
             int(blocks_per_cu), int(forward_tile), ctypes.c_void_p(ws.data_ptr()), ws_bytes,
             ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
    return outs


def _device_geometry(items, geometries, dev):
    """Caller-held device arrays of each problem (objects with float32 device tensors P
    [.., n, 3, 4], vo and vs [.., 3], step and mult [.., 1], and g, the geometry), e.g.
    the server's FDK plans; checked against the geometries and the device."""
    if len(items) != len(geometries):
        raise ValueError(f"{len(items)} device geometries for {len(geometries)} problems")
    for k, (it, g) in enumerate(zip(items, geometries)):
        if it.g is not g:
            raise ValueError(f"device geometry {k} belongs to another geometry")
        for name, n in (("P", g.n_views * 12), ("vo", 3), ("vs", 3), ("step", 1), ("mult", 1)):
            t = getattr(it, name)
            if (not t.is_cuda or t.device != dev or t.dtype != torch.float32 or not t.is_contiguous()
                    or t.numel() != n):
                raise ValueError(f"device geometry {k}: {name} must be a contiguous float32 tensor of {n} "
                                 f"values on {dev}")
    return list(items)


def cone_backprojection3d_batched(sinograms: Sequence[torch.Tensor], geometries: Sequence[geo.ConeGeometry],
                                  hardware_interp: bool = False, z_per_lane: int = 0,
                                  blocks_per_cu: int = 0, device_geometry=None) -> List[torch.Tensor]:
    """sinograms[k] [n_k, H_d,k, W_d,k] -> volumes [D_k, H_k, W_k] float32 (scaled by projection_multiplier).
    device_geometry: optional per-problem device arrays already resident (P, vo, vs, step, mult;
    e.g. FDK plans, whose mult is the FDK scale) used instead of uploading the geometries'."""
    _validate(sinograms, geometries, False)
    dev = sinograms[0].device
    dt = sinograms[0].dtype
    if any(s.dtype != dt for s in sinograms):
        raise TypeError("all sinograms of a batched launch share one storage dtype")
    sins = [s.contiguous() for s in sinograms]
    geos = (_device_geometry(device_geometry, geometries, dev) if device_geometry is not None
            else [_Geo(g, dev) for g in geometries])
    outs = [torch.empty(tuple(g.volume_shape), dtype=torch.float32, device=dev) for g in geometries]
    arr = _problems(sins, outs, geos)
    L = nat.lib()
    ws_bytes = L.pyronn_cone_backprojection3d_batched_workspace_size(arr, len(sins), _DT[dt])
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    nat.call("pyronn_cone_backprojection3d_batched", arr, len(sins), _DT[dt], int(bool(hardware_interp)),
             int(z_per_lane), int(blocks_per_cu), ctypes.c_void_p(ws.data_ptr()), ws_bytes,
             ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
    return outs
