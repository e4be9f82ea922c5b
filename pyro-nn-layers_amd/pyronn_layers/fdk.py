"""Filtered backprojection: FDK (cone, circular orbit) and FBP (fan / parallel).

Not part of PYRO-NN-Layers itself: the reference's backprojectors expect
weighted, filtered projections and take the FDK normalisation as
`projection_multiplier` (cc/kernels/cone_backprojector_3D_CudaKernel.cu.cc:117);
weighting and filtering lived in the companion pyronn package.  Here they are
formulated after Kak & Slaney, ch. 3, for full (2 pi) orbits:

  cone  f(x) = (pi / n) SID^2 sum_n Q_n(u_n, v_n) / w_n^2,  Q = cosine-weighted
        projection filtered along detector rows with the Ram-Lak kernel at the
        isocentre sample spacing tau = du SID / SDD; w_n is the depth along the
        central ray, exactly the reference backprojector's 1/w^2 weight, so the
        plan passes projection_multiplier = pi SID^2 / n.
  fan   the same with one detector row.  The reference fan backprojector
        (fan_backprojector_2D_CudaKernel.cu.cc:56-64) weights by 1/|x - s|^2 and
        scales by SID SDD pi / n; |x - s| = depth / cos(gamma) with gamma fixed by
        the detector position, so the filter output is multiplied by
        (SDD^2 + s^2) / SDD^2 * SID / SDD to make the sum the textbook one.
  par   f(x) = (pi / n) sum_n Q_n(x . theta_n) (the reference's pi / n scaling),
        Q = R filtered at tau = detector spacing.

Plans keep all geometry and filter state resident on the device (projection
matrices, cosine weights, the ramp spectrum and FFT twiddles), so one
reconstruction is: one fused weighting + FFT ramp filter + storage conversion
kernel and one backprojection, all on the current stream and capturable in a HIP
graph (see `capture`).  Rows wider than 4096 samples take the Toeplitz-matrix
GEMM filter (weighting kernel + rocBLAS SGEMM + conversion) instead.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import geometry as geo
from . import load as _load_ops
from . import native as nat
from .projectors import Capturable

_DT = {torch.float32: nat.F32, torch.bfloat16: nat.BF16, torch.float16: nat.F16}


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _cuda_device(device) -> torch.device:
    """The plan's device with its index filled in, so that tensors can be compared against it."""
    d = torch.device(device)
    if d.type == "cuda" and d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


class _Plan(Capturable):
    storage: torch.dtype
    device: torch.device
    # "fft": the fused LDS FFT filter (detector rows up to 4096 samples); "gemm": the
    # Toeplitz-matrix product in rocBLAS (any width).  Both compute the same convolution.
    filter_impl: str = "fft"

    def _init_filter(self, det_w: int, tau: float, column_weights: torch.Tensor | None = None):
        self.tau = tau
        self._det_w = det_w
        self._colw = column_weights
        self.ramp = None
        self.spectrum = None
        n = nat.lib().pyronn_fdk_ramp_spectrum_size(det_w)
        if n > 0:
            self.spectrum = torch.empty(n, dtype=torch.float32, device=self.device)
            nat.call("pyronn_fdk_ramp_spectrum", det_w, tau, _p(column_weights) if column_weights is not None
                     else None, _p(self.spectrum), _stream())
        else:
            self.filter_impl = "gemm"
            self.ramp_matrix()

    def footprint(self) -> int:
        """Device bytes of one in-flight request: its input, the filtered copy, the GEMM
        filter's workspace (two float32 copies) if that filter is used, the detector cells
        of the cone backprojection (four values per pixel) and the float32 output."""
        nin = int(np.prod(self.input_shape))
        nout = int(np.prod(self.output_shape))
        e = torch.tensor([], dtype=self.storage).element_size()
        ws = 8 if self.filter_impl == "gemm" else 0
        cells = 4 * e if isinstance(self, ConeFDK) else 0
        return nin * (e + e + ws + cells) + nout * 4

    def ramp_matrix(self) -> torch.Tensor:
        """The Toeplitz ramp matrix of the GEMM filter (built on first use)."""
        if self.ramp is None:
            W = self._det_w
            self.ramp = torch.empty((W, W), dtype=torch.float32, device=self.device)
            nat.call("pyronn_fdk_ramp_matrix", W, self.tau, _p(self._colw) if self._colw is not None else None,
                     _p(self.ramp), _stream())
        return self.ramp

    def _filter(self, sino: torch.Tensor, images: int, det_h: int, det_w: int) -> torch.Tensor:
        # the cosine weights (det_h x det_w) and the filter table were built for the plan's
        # detector: a sinogram of any other row length or row count would read past them
        if det_w != self._det_w or (self.weights is not None and self.weights.numel() != det_h * det_w):
            raise ValueError(f"sinogram rows of {det_h} x {det_w} samples; the plan was built for "
                             f"{tuple(self.weights.shape) if self.weights is not None else self._det_w}")
        if not sino.is_cuda or sino.device != self.device:
            raise ValueError(f"sinogram on {sino.device}, the plan on {self.device}")
        sino = sino.contiguous()
        code = _DT.get(sino.dtype)
        if code is None:
            raise TypeError(f"sinogram dtype {sino.dtype} not supported")
        out = torch.empty(sino.shape, dtype=self.storage, device=self.device)
        w = _p(self.weights) if self.weights is not None else None
        if self.filter_impl == "fft":
            nat.call("pyronn_fdk_filter_fft", _p(sino), code, images, det_h, det_w, w, _p(self.spectrum), _p(out),
                     _DT[self.storage], _stream())
            return out
        ramp = self.ramp_matrix()
        # Per-call workspace from torch's stream-aware caching allocator (graph-pool
        # memory under capture), so concurrent calls on different streams never share it.
        ws_bytes = nat.lib().pyronn_fdk_filter_workspace_size(images, det_h, det_w)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=self.device)
        nat.call("pyronn_fdk_filter", _p(sino), code, images, det_h, det_w, w, _p(ramp), _p(out),
                 _DT[self.storage], _p(ws), ws.numel(), _stream())
        return out


class ConeFDK(_Plan):
    """FDK for a full circular cone-beam orbit described by a ConeGeometry."""

    def __init__(self, g: geo.ConeGeometry, device="cuda", storage: torch.dtype = torch.float32):
        _load_ops()
        self.g = g
        self.device = _cuda_device(device)
        self.storage = storage
        n = g.n_views
        Hd, Wd = g.detector_shape
        dv, du = g.detector_spacing
        dev = self.device
        self.P = torch.tensor(g.projection_matrices()[None], dtype=torch.float32, device=dev)
        self.vo = torch.tensor(np.asarray(g.volume_origin)[None], dtype=torch.float32, device=dev)
        self.vs = torch.tensor(np.asarray(g.volume_spacing)[None], dtype=torch.float32, device=dev)
        self.step = torch.full((1, 1), g.step_size, dtype=torch.float32, device=dev)
        dbeta = g.angular_range / n
        self.mult_value = 0.5 * dbeta * g.sid ** 2
        self.mult = torch.full((1, 1), self.mult_value, dtype=torch.float32, device=dev)
        self.weights = torch.empty((Hd, Wd), dtype=torch.float32, device=dev)
        nat.call("pyronn_fdk_cosine_weights", Hd, Wd, du, dv, (Wd - 1) / 2.0, (Hd - 1) / 2.0, g.sdd,
                 _p(self.weights), _stream())
        self._init_filter(Wd, du * g.sid / g.sdd)
        self.input_shape = (1, n, Hd, Wd)
        self.output_shape = (1,) + tuple(g.volume_shape)

    def filter(self, sino: torch.Tensor) -> torch.Tensor:
        B, n, Hd, Wd = sino.shape
        return self._filter(sino, B * n, Hd, Wd)

    def __call__(self, sino: torch.Tensor) -> torch.Tensor:
        """sino [B, n, H_d, W_d] (float32 / bfloat16 / float16) -> volume [B, D, H, W] float32."""
        if sino.dim() == 3:
            sino = sino[None]
        q = self.filter(sino)
        return torch.ops.pyronn_layers.cone_backprojection3d(q, list(self.g.volume_shape), self.vo, self.vs,
                                                             self.P, self.step, self.mult, False)

    def backproject_views(self, q: torch.Tensor, view0: int, mode: int, out: torch.Tensor, z=None,
                          ws: torch.Tensor | None = None) -> torch.Tensor:
        """Backproject filtered views q [1, k, H_d, W_d] (views view0 .. view0 + k - 1 of the
        orbit) into the running sums in out [1, D, H, W]; mode as
        pyronn_cone_backprojection3d_views (bit 0 continue, bit 1 apply the FDK scale, bit 2
        reuse the detector cells already in `ws`).  z = (z0, z1) limits the slices computed.
        Returns the workspace (holding q's detector cells) for a following call with bit 2."""
        _, k, Hd, Wd = q.shape
        D, H, W = self.g.volume_shape
        if q.dim() != 4 or q.shape[0] != 1 or (Hd, Wd) != tuple(self.g.detector_shape):
            raise ValueError(f"filtered views {tuple(q.shape)}; expected [1, k, {self.g.detector_shape[0]}, "
                             f"{self.g.detector_shape[1]}]")
        if not 0 <= view0 or view0 + k > self.g.n_views:
            raise ValueError(f"views {view0} .. {view0 + k - 1} outside the orbit's {self.g.n_views} views")
        if tuple(out.shape) != (1, D, H, W) or out.dtype != torch.float32 or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous float32 [1, {D}, {H}, {W}] tensor")
        z0, z1 = z if z is not None else (0, D)
        code = _DT[q.dtype]
        ws_bytes = nat.lib().pyronn_cone_backprojection3d_workspace_size(1, k, Hd, Wd, code)
        if ws is None:
            if mode & 4:
                raise ValueError("mode bit 2 needs the workspace of the previous call")
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=self.device)
        P = self.P[:, view0:view0 + k]
        nat.call("pyronn_cone_backprojection3d_views", _p(q), code, 1, k, Hd, Wd, _p(self.vo), _p(self.vs), _p(P),
                 _p(self.mult), D, H, W, 0, mode, z0, z1, _p(out), _p(ws), ws.numel(), _stream())
        return ws

    def chunk_bounds(self, chunks: int):
        n = self.g.n_views
        edges = [round(i * n / chunks) for i in range(chunks + 1)]
        return [(a, b) for a, b in zip(edges[:-1], edges[1:]) if b > a]

    def run_chunked(self, host_sino: torch.Tensor, out: torch.Tensor, bufs, copy_stream, chunks: int = 4,
                    host_out: torch.Tensor | None = None, out_slabs: int = 4):
        """FDK of a sinogram in pinned host memory [1, n, H_d, W_d] into out [1, D, H, W].
        The views are uploaded in `chunks` ranges on copy_stream, into the two device
        buffers `bufs` (each [1, >= max chunk, H_d, W_d] of the sinogram's dtype) in turn,
        while the current stream weights, filters and backprojects the previous range.
        The result equals self(device_sinogram) bit for bit: the FFT filter treats each
        view alone and the backprojection continues each voxel's running sum across
        ranges.  With host_out (pinned, shaped like out) the last range is backprojected in
        `out_slabs` z slabs, each downloaded on copy_stream while the next is computed.
        Returns after queueing; synchronize the current stream before reading out or
        host_out."""
        cur = torch.cuda.current_stream()
        ranges = self.chunk_bounds(chunks)
        free = [torch.cuda.Event(), torch.cuda.Event()]
        for e in free:
            e.record(cur)
        D = self.g.volume_shape[0]
        for i, (a, b) in enumerate(ranges):
            buf = bufs[i % 2][:, :b - a]
            arrived = torch.cuda.Event()
            with torch.cuda.stream(copy_stream):
                copy_stream.wait_event(free[i % 2])          # the range that used this buffer is done
                buf.copy_(host_sino[:, a:b], non_blocking=True)
                arrived.record(copy_stream)
            cur.wait_event(arrived)
            q = self.filter(buf)
            last = i == len(ranges) - 1
            mode = (1 if i > 0 else 0) | (2 if last else 0)
            if not (last and host_out is not None):
                self.backproject_views(q, a, mode, out)
            else:
                edges = [round(j * D / out_slabs) for j in range(out_slabs + 1)]
                ws = None
                for z0, z1 in zip(edges[:-1], edges[1:]):
                    if z1 <= z0:
                        continue
                    ws = self.backproject_views(q, a, mode | (4 if ws is not None else 0), out, (z0, z1), ws)
                    done = torch.cuda.Event()
                    done.record(cur)
                    with torch.cuda.stream(copy_stream):
                        copy_stream.wait_event(done)
                        host_out[:, z0:z1].copy_(out[:, z0:z1], non_blocking=True)
                downloaded = torch.cuda.Event()
                downloaded.record(copy_stream)
                cur.wait_event(downloaded)
            free[i % 2] = torch.cuda.Event()
            free[i % 2].record(cur)


class FanFBP(_Plan):
    """FBP for a full fan-beam orbit (equispaced flat detector) given a Geometry2D with sid/sdd."""

    def __init__(self, g: geo.Geometry2D, device="cuda", storage: torch.dtype = torch.float32):
        _load_ops()
        self.g = g
        self.device = _cuda_device(device)
        self.storage = storage
        dev = self.device
        det, ds = g.detector_shape, g.detector_spacing
        self.rays = torch.tensor(geo.fan_central_ray_vectors(g.n_views, g.angular_range)[None],
                                 dtype=torch.float32, device=dev)
        self.vo = torch.tensor(np.asarray(g.volume_origin)[None], dtype=torch.float32, device=dev)
        self.vs = torch.tensor(np.asarray(g.volume_spacing)[None], dtype=torch.float32, device=dev)
        self.do = torch.full((1, 1), g.detector_origin, dtype=torch.float32, device=dev)
        self.ds = torch.full((1, 1), ds, dtype=torch.float32, device=dev)
        self.sid = torch.full((1, 1), g.sid, dtype=torch.float32, device=dev)
        self.sdd = torch.full((1, 1), g.sdd, dtype=torch.float32, device=dev)
        u0 = -g.detector_origin / ds
        self.weights = torch.empty((1, det), dtype=torch.float32, device=dev)
        nat.call("pyronn_fdk_cosine_weights", 1, det, ds, 0.0, u0, 0.0, g.sdd, _p(self.weights), _stream())
        s = g.detector_origin + ds * np.arange(det)
        # 1/cos^2(gamma) undoes the reference's Euclidean 1/|x-s|^2 weight, SID/SDD its SDD factor;
        # full orbit: (1/2) dbeta = pi / n, which the reference already applies.
        post = ((g.sdd ** 2 + s ** 2) / g.sdd ** 2) * (g.sid / g.sdd) * (g.angular_range / (2 * math.pi))
        colw = torch.tensor(post, dtype=torch.float32, device=dev)
        self._init_filter(det, ds * g.sid / g.sdd, colw)
        self.input_shape = (1, g.n_views, det)
        self.output_shape = (1,) + tuple(g.volume_shape)
        torch.cuda.current_stream().synchronize()

    def filter(self, sino: torch.Tensor) -> torch.Tensor:
        """sino [B, n, det] -> weighted, ramp-filtered rows (plan storage dtype)."""
        if sino.dim() == 2:
            sino = sino[None]
        B, n, det = sino.shape
        return self._filter(sino, B * n, 1, det)

    def __call__(self, sino: torch.Tensor) -> torch.Tensor:
        """sino [B, n, det] -> image [B, H, W] float32."""
        q = self.filter(sino)
        return torch.ops.pyronn_layers.fan_backprojection2d(q, list(self.g.volume_shape), self.vo, self.do,
                                                            self.vs, self.ds, self.sid, self.sdd, self.rays)


class ParallelFBP(_Plan):
    """FBP for parallel-beam views over pi (Geometry2D)."""

    def __init__(self, g: geo.Geometry2D, device="cuda", storage: torch.dtype = torch.float32):
        _load_ops()
        self.g = g
        self.device = _cuda_device(device)
        self.storage = storage
        dev = self.device
        det, ds = g.detector_shape, g.detector_spacing
        self.rays = torch.tensor(geo.parallel_ray_vectors(g.n_views, g.angular_range)[None], dtype=torch.float32,
                                 device=dev)
        self.vo = torch.tensor(np.asarray(g.volume_origin)[None], dtype=torch.float32, device=dev)
        self.vs = torch.tensor(np.asarray(g.volume_spacing)[None], dtype=torch.float32, device=dev)
        self.do = torch.full((1, 1), g.detector_origin, dtype=torch.float32, device=dev)
        self.ds = torch.full((1, 1), ds, dtype=torch.float32, device=dev)
        self.weights = None
        # The reference's pi / n scaling is the FBP weight for views over pi and,
        # since every direction then appears twice, also for views over 2 pi.
        self._init_filter(det, ds)
        self.input_shape = (1, g.n_views, det)
        self.output_shape = (1,) + tuple(g.volume_shape)
        torch.cuda.current_stream().synchronize()

    def filter(self, sino: torch.Tensor) -> torch.Tensor:
        """sino [B, n, det] -> ramp-filtered rows (plan storage dtype)."""
        if sino.dim() == 2:
            sino = sino[None]
        B, n, det = sino.shape
        return self._filter(sino, B * n, 1, det)

    def __call__(self, sino: torch.Tensor) -> torch.Tensor:
        q = self.filter(sino)
        return torch.ops.pyronn_layers.parallel_backprojection2d(q, list(self.g.volume_shape), self.vo, self.do,
                                                                 self.vs, self.ds, self.rays)
