"""pyronn_layers on MI355X: the six PYRO-NN-Layers reconstruction operators.

Same names, argument order and meaning as the reference's generated wrappers
(python/ops/pyronn_layers_ops.py:27-29 exposes the REGISTER_OP names of
cc/ops/*_OPKernel.cc in snake case):

    parallel_projection2d(volume, projection_shape, volume_origin, detector_origin,
                          volume_spacing, detector_spacing, ray_vectors)
    parallel_backprojection2d(sinogram, volume_shape, volume_origin, detector_origin,
                              volume_spacing, detector_spacing, ray_vectors)
    fan_projection2d(volume, projection_shape, volume_origin, detector_origin,
                     volume_spacing, detector_spacing, source_2_isocenter_distance,
                     source_2_detector_distance, central_ray_vectors)
    fan_backprojection2d(sinogram, volume_shape, ...same...)
    cone_projection3d(volume, projection_shape, volume_origin, volume_spacing,
                      projection_matrices, step_size, projection_multiplier,
                      hardware_interp=False)
    cone_backprojection3d(sinogram, volume_shape, ...same..., hardware_interp=False)

Inputs are torch tensors on the GPU with a leading batch dimension.  Each
projection is differentiable and its gradient is the paired backprojection of
the incoming gradient with the same geometry (and vice versa) - the pairing the
companion pyronn package registers for the TF ops.  The pair is unmatched (not
an exact adjoint), exactly as in the reference.

The kernels run in libpyronn_mi.so (HIP, gfx950) through the op library
libpyronn_torch_ops.so; there is no CPU path.
"""
from __future__ import annotations

import os
from typing import Sequence, Union

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libpyronn_mi.so")
OPS_LIB = os.path.join(LIB_DIR, "libpyronn_torch_ops.so")

_loaded = False


def load() -> None:
    """Load the op library; raises if it has not been built."""
    global _loaded
    if _loaded:
        return
    if not os.path.exists(OPS_LIB) or not os.path.exists(HIP_LIB):
        raise RuntimeError(
            f"pyronn_layers: native libraries missing ({OPS_LIB}); build them with "
            "`python pyro-nn-layers_amd/build.py`")
    torch.ops.load_library(OPS_LIB)
    _loaded = True


def _shape(s: Union[torch.Tensor, Sequence[int]]) -> list:
    if isinstance(s, torch.Tensor):
        return [int(v) for v in s.detach().cpu().reshape(-1).tolist()]
    return [int(v) for v in s]


def _ops():
    load()
    return torch.ops.pyronn_layers


# ----------------------------------------------------------------------------- autograd pairs

class _ParallelProjection2D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, volume, shape, vo, dor, vs, ds, rays, step):
        ctx.save_for_backward(vo, dor, vs, ds, rays)
        ctx.vshape = list(volume.shape[1:])
        ctx.vdtype = volume.dtype
        return _ops().parallel_projection2d(volume, shape, vo, dor, vs, ds, rays, step)

    @staticmethod
    def backward(ctx, grad):
        vo, dor, vs, ds, rays = ctx.saved_tensors
        g = _ops().parallel_backprojection2d(grad.contiguous(), ctx.vshape, vo, dor, vs, ds, rays)
        return g.to(ctx.vdtype), None, None, None, None, None, None, None


class _ParallelBackprojection2D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, shape, vo, dor, vs, ds, rays):
        ctx.save_for_backward(vo, dor, vs, ds, rays)
        ctx.pshape = list(sino.shape[1:])
        ctx.sdtype = sino.dtype
        return _ops().parallel_backprojection2d(sino, shape, vo, dor, vs, ds, rays)

    @staticmethod
    def backward(ctx, grad):
        vo, dor, vs, ds, rays = ctx.saved_tensors
        g = _ops().parallel_projection2d(grad.contiguous(), ctx.pshape, vo, dor, vs, ds, rays, 0.2)
        return g.to(ctx.sdtype), None, None, None, None, None, None


class _FanProjection2D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, volume, shape, vo, dor, vs, ds, sid, sdd, rays, step):
        ctx.save_for_backward(vo, dor, vs, ds, sid, sdd, rays)
        ctx.vshape = list(volume.shape[1:])
        ctx.vdtype = volume.dtype
        return _ops().fan_projection2d(volume, shape, vo, dor, vs, ds, sid, sdd, rays, step)

    @staticmethod
    def backward(ctx, grad):
        vo, dor, vs, ds, sid, sdd, rays = ctx.saved_tensors
        g = _ops().fan_backprojection2d(grad.contiguous(), ctx.vshape, vo, dor, vs, ds, sid, sdd, rays)
        return (g.to(ctx.vdtype),) + (None,) * 9


class _FanBackprojection2D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, shape, vo, dor, vs, ds, sid, sdd, rays):
        ctx.save_for_backward(vo, dor, vs, ds, sid, sdd, rays)
        ctx.pshape = list(sino.shape[1:])
        ctx.sdtype = sino.dtype
        return _ops().fan_backprojection2d(sino, shape, vo, dor, vs, ds, sid, sdd, rays)

    @staticmethod
    def backward(ctx, grad):
        vo, dor, vs, ds, sid, sdd, rays = ctx.saved_tensors
        g = _ops().fan_projection2d(grad.contiguous(), ctx.pshape, vo, dor, vs, ds, sid, sdd, rays, 0.2)
        return (g.to(ctx.sdtype),) + (None,) * 8


class _ConeProjection3D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, volume, shape, vo, vs, P, step, mult, hw):
        ctx.save_for_backward(vo, vs, P, step, mult)
        ctx.vshape = list(volume.shape[1:])
        ctx.vdtype = volume.dtype
        ctx.hw = hw
        return _ops().cone_projection3d(volume, shape, vo, vs, P, step, mult, hw)

    @staticmethod
    def backward(ctx, grad):
        vo, vs, P, step, mult = ctx.saved_tensors
        g = _ops().cone_backprojection3d(grad.contiguous(), ctx.vshape, vo, vs, P, step, mult, ctx.hw)
        return (g.to(ctx.vdtype),) + (None,) * 7


class _ConeBackprojection3D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, shape, vo, vs, P, step, mult, hw):
        ctx.save_for_backward(vo, vs, P, step, mult)
        ctx.pshape = list(sino.shape[1:])
        ctx.sdtype = sino.dtype
        ctx.hw = hw
        return _ops().cone_backprojection3d(sino, shape, vo, vs, P, step, mult, hw)

    @staticmethod
    def backward(ctx, grad):
        vo, vs, P, step, mult = ctx.saved_tensors
        g = _ops().cone_projection3d(grad.contiguous(), ctx.pshape, vo, vs, P, step, mult, ctx.hw)
        return (g.to(ctx.sdtype),) + (None,) * 7


# ----------------------------------------------------------------------------- public ops

def parallel_projection2d(volume, projection_shape, volume_origin, detector_origin, volume_spacing,
                          detector_spacing, ray_vectors, step_size: float = 0.2, name=None):
    """ParallelProjection2D (cc/ops/par_projector_2D_OPKernel.cc:30-56) -> [B, n, det]."""
    return _ParallelProjection2D.apply(volume, _shape(projection_shape), volume_origin, detector_origin,
                                       volume_spacing, detector_spacing, ray_vectors, float(step_size))


def parallel_backprojection2d(sinogram, volume_shape, volume_origin, detector_origin, volume_spacing,
                              detector_spacing, ray_vectors, name=None):
    """ParallelBackprojection2D (cc/ops/par_backprojector_2D_OPKernel.cc:26-51) -> [B, H, W]."""
    return _ParallelBackprojection2D.apply(sinogram, _shape(volume_shape), volume_origin, detector_origin,
                                           volume_spacing, detector_spacing, ray_vectors)


def fan_projection2d(volume, projection_shape, volume_origin, detector_origin, volume_spacing,
                     detector_spacing, source_2_isocenter_distance, source_2_detector_distance,
                     central_ray_vectors, step_size: float = 0.2, name=None):
    """FanProjection2D (cc/ops/fan_projector_2D_OPKernel.cc:26-53) -> [B, n, det]."""
    return _FanProjection2D.apply(volume, _shape(projection_shape), volume_origin, detector_origin,
                                  volume_spacing, detector_spacing, source_2_isocenter_distance,
                                  source_2_detector_distance, central_ray_vectors, float(step_size))


def fan_backprojection2d(sinogram, volume_shape, volume_origin, detector_origin, volume_spacing,
                         detector_spacing, source_2_isocenter_distance, source_2_detector_distance,
                         central_ray_vectors, name=None):
    """FanBackprojection2D (cc/ops/fan_backprojector_2D_OPKernel.cc:26-53) -> [B, H, W]."""
    return _FanBackprojection2D.apply(sinogram, _shape(volume_shape), volume_origin, detector_origin,
                                      volume_spacing, detector_spacing, source_2_isocenter_distance,
                                      source_2_detector_distance, central_ray_vectors)


def cone_projection3d(volume, projection_shape, volume_origin, volume_spacing, projection_matrices,
                      step_size, projection_multiplier, hardware_interp: bool = False, name=None):
    """ConeProjection3D (cc/ops/cone_projector_3D_OPKernel.cc:34-60) -> [B, n, H_d, W_d]."""
    return _ConeProjection3D.apply(volume, _shape(projection_shape), volume_origin, volume_spacing,
                                   projection_matrices, step_size, projection_multiplier,
                                   bool(hardware_interp))


def cone_backprojection3d(sinogram, volume_shape, volume_origin, volume_spacing, projection_matrices,
                          step_size, projection_multiplier, hardware_interp: bool = False, name=None):
    """ConeBackprojection3D (cc/ops/cone_backprojector_3D_OPKernel.cc:27-53) -> [B, D, H, W]."""
    return _ConeBackprojection3D.apply(sinogram, _shape(volume_shape), volume_origin, volume_spacing,
                                       projection_matrices, step_size, projection_multiplier,
                                       bool(hardware_interp))


__all__ = [
    "parallel_projection2d", "parallel_backprojection2d", "fan_projection2d", "fan_backprojection2d",
    "cone_projection3d", "cone_backprojection3d", "load",
]
