"""Projection plans: forward projections with their geometry resident on the device.

A plan holds one geometry's arrays as device tensors (projection matrices, origin,
spacing, step, or the 2D ray vectors and detector scalars) so that a request is only
its volume: `plan(volume)` runs the reference operator (cone_projection3d,
fan_projection2d, parallel_projection2d; cc/ops/cone_projector_3D_OPKernel.cc:185-248,
cc/ops/fan_projector_2D_OPKernel.cc:66-128, cc/ops/par_projector_2D_OPKernel.cc:67-120)
on the current stream, and `capture()` records it in a HIP graph.  These are what the
server runs for projection requests, beside the FDK / FBP plans of fdk.py; every plan
exposes the same interface:

    storage         dtype of the request input
    input_shape     shape of one request's input (batch 1)
    output_shape    shape of its float32 output
    footprint()     device bytes one in-flight request needs (buffers + workspace)
    __call__(x)     run on the current stream
    capture(x)      (graph, static_input, static_output)
"""
from __future__ import annotations

import numpy as np
import torch

from . import geometry as geo
from . import load as _load_ops


def _cuda_device(device) -> torch.device:
    d = torch.device(device)
    if d.type == "cuda" and d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


class Capturable:
    """HIP-graph capture of a plan's __call__ (shared by projection and FDK plans)."""

    def capture(self, example: torch.Tensor, warmup: int = 2):
        """Capture the plan for inputs shaped like `example` in a HIP graph.  Returns
        (graph, static_input, static_output): copy a request into static_input, replay
        the graph, read static_output."""
        static_in = example.clone()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self(static_in)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            static_out = self(static_in)
        return g, static_in, static_out

    def _check_input(self, x: torch.Tensor) -> torch.Tensor:
        if x.dim() == len(self.input_shape) - 1:
            x = x[None]
        if tuple(x.shape) != tuple(self.input_shape):
            raise ValueError(f"input of shape {tuple(x.shape)}; the plan takes {tuple(self.input_shape)}")
        if not x.is_cuda or x.device != self.device:
            raise ValueError(f"input on {x.device}, the plan on {self.device}")
        if x.dtype != self.storage:
            raise TypeError(f"input dtype {x.dtype}; the plan was built for {self.storage}")
        return x


class ConeProjector(Capturable):
    """cone_projection3d of one volume [1, D, H, W] for a ConeGeometry."""

    def __init__(self, g: geo.ConeGeometry, device="cuda", storage: torch.dtype = torch.float32,
                 hardware_interp: bool = False):
        _load_ops()
        self.g = g
        self.device = _cuda_device(device)
        self.storage = storage
        self.hardware_interp = bool(hardware_interp)
        dev = self.device
        f32 = dict(dtype=torch.float32, device=dev)
        self.P = torch.tensor(g.projection_matrices()[None], **f32)
        self.vo = torch.tensor(np.asarray(g.volume_origin)[None], **f32)
        self.vs = torch.tensor(np.asarray(g.volume_spacing)[None], **f32)
        self.step = torch.full((1, 1), g.step_size, **f32)
        self.mult = torch.full((1, 1), g.projection_multiplier, **f32)
        self.input_shape = (1,) + tuple(g.volume_shape)
        self.output_shape = (1,) + tuple(g.projection_shape)

    def footprint(self) -> int:
        D, H, W = self.g.volume_shape
        e = torch.tensor([], dtype=self.storage).element_size()
        cells = (W + 2) * (H + 2) * (D + 3) * 4 * e          # the forward's cell workspace
        return D * H * W * e + cells + int(np.prod(self.output_shape)) * 4 + self.g.n_views * 48

    def __call__(self, volume: torch.Tensor) -> torch.Tensor:
        volume = self._check_input(volume)
        return torch.ops.pyronn_layers.cone_projection3d(volume, list(self.g.projection_shape), self.vo, self.vs,
                                                         self.P, self.step, self.mult, self.hardware_interp)


class FanProjector(Capturable):
    """fan_projection2d of one image [1, H, W] for a Geometry2D with sid / sdd."""

    def __init__(self, g: geo.Geometry2D, device="cuda", storage: torch.dtype = torch.float32,
                 step_factor: float | None = None):
        _load_ops()
        self.g = g
        self.device = _cuda_device(device)
        self.storage = storage
        self.step_factor = float(g.step_size if step_factor is None else step_factor)
        dev = self.device
        f32 = dict(dtype=torch.float32, device=dev)
        self.rays = torch.tensor(geo.fan_central_ray_vectors(g.n_views, g.angular_range)[None], **f32)
        self.vo = torch.tensor(np.asarray(g.volume_origin)[None], **f32)
        self.vs = torch.tensor(np.asarray(g.volume_spacing)[None], **f32)
        self.do = torch.full((1, 1), g.detector_origin, **f32)
        self.ds = torch.full((1, 1), g.detector_spacing, **f32)
        self.sid = torch.full((1, 1), g.sid, **f32)
        self.sdd = torch.full((1, 1), g.sdd, **f32)
        self.input_shape = (1,) + tuple(g.volume_shape)
        self.output_shape = (1,) + tuple(g.projection_shape)

    def footprint(self) -> int:
        e = torch.tensor([], dtype=self.storage).element_size()
        return int(np.prod(self.input_shape)) * e + int(np.prod(self.output_shape)) * 4 + self.g.n_views * 8

    def __call__(self, image: torch.Tensor) -> torch.Tensor:
        image = self._check_input(image)
        return torch.ops.pyronn_layers.fan_projection2d(image, list(self.g.projection_shape), self.vo, self.do,
                                                        self.vs, self.ds, self.sid, self.sdd, self.rays,
                                                        self.step_factor)


class ParallelProjector(Capturable):
    """parallel_projection2d of one image [1, H, W] for a Geometry2D."""

    def __init__(self, g: geo.Geometry2D, device="cuda", storage: torch.dtype = torch.float32,
                 step_factor: float | None = None):
        _load_ops()
        self.g = g
        self.device = _cuda_device(device)
        self.storage = storage
        self.step_factor = float(g.step_size if step_factor is None else step_factor)
        dev = self.device
        f32 = dict(dtype=torch.float32, device=dev)
        self.rays = torch.tensor(geo.parallel_ray_vectors(g.n_views, g.angular_range)[None], **f32)
        self.vo = torch.tensor(np.asarray(g.volume_origin)[None], **f32)
        self.vs = torch.tensor(np.asarray(g.volume_spacing)[None], **f32)
        self.do = torch.full((1, 1), g.detector_origin, **f32)
        self.ds = torch.full((1, 1), g.detector_spacing, **f32)
        self.input_shape = (1,) + tuple(g.volume_shape)
        self.output_shape = (1,) + tuple(g.projection_shape)

    def footprint(self) -> int:
        e = torch.tensor([], dtype=self.storage).element_size()
        return int(np.prod(self.input_shape)) * e + int(np.prod(self.output_shape)) * 4 + self.g.n_views * 8

    def __call__(self, image: torch.Tensor) -> torch.Tensor:
        image = self._check_input(image)
        return torch.ops.pyronn_layers.parallel_projection2d(image, list(self.g.projection_shape), self.vo,
                                                             self.do, self.vs, self.ds, self.rays,
                                                             self.step_factor)
