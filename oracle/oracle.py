"""ctypes front end for the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py.  The product library never imports it.

Array layouts follow the reference ops (cc/ops/*_OPKernel.cc): volumes are
[D,H,W] / [H,W], sinograms [n,H_d,W_d] / [n,det], origin and spacing in
(z,y,x) / (y,x) order, projection matrices [n,3,4] row-major, 2D ray vectors
[n,2] as (x,y).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = ctypes.POINTER(ctypes.c_longlong)
_vp = ctypes.c_void_p


def build() -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        L.orc_cone_setup.restype = ctypes.c_int
        L.orc_cone_setup.argtypes = [_f32p, ctypes.c_int, _f32p, _f32p, _vp, _vp, _vp, _vp]
        L.orc_cone_fwd_f32.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f32p,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _f32p, ctypes.c_float, ctypes.c_int, ctypes.c_int, _f32p, _vp, _i64p]
        L.orc_cone_fwd_f64.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, _f64p,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _f64p, ctypes.c_double, ctypes.c_int, _f64p, _i64p]
        L.orc_cone_bp_f32.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, _f32p, _f32p, ctypes.c_float, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _vp]
        L.orc_cone_bp_f64.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, _f64p, _f64p, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, _f64p]
        L.orc_proj2d_f32.argtypes = [ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_int,
                                     ctypes.c_int, _f32p, _f32p, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_int, _f32p, _vp, _i64p]
        L.orc_proj2d_f64.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int,
                                     ctypes.c_int, _f64p, _f64p, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, _f64p, _i64p]
        L.orc_back2d_f32.argtypes = [ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_int,
                                     ctypes.c_int, _f32p, _f32p, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_int, _f32p, _vp]
        L.orc_back2d_f64.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int,
                                     ctypes.c_int, _f64p, _f64p, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, _f64p]
        _lib = L
    return _lib


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------- cone


def cone_setup(P, spacing_zyx, origin_zyx):
    """FP64 closed-form per-view setup -> (inv_ar f32[n,9], src f32[n,3], inv_ar f64, src f64, status)."""
    P = _c(P, np.float32).reshape(-1, 12)
    n = P.shape[0]
    inv = np.empty((n, 9), np.float32)
    src = np.empty((n, 3), np.float32)
    inv64 = np.empty((n, 9), np.float64)
    src64 = np.empty((n, 3), np.float64)
    st = lib().orc_cone_setup(P, n, _c(spacing_zyx, np.float32), _c(origin_zyx, np.float32),
                              _ptr(inv), _ptr(src), _ptr(inv64), _ptr(src64))
    return inv, src, inv64, src64, st


def cone_forward(vol, inv_ar, src, det_shape, spacing_zyx, step, hw=False, qbits=0, views=None,
                 want_bound=False):
    """f32-faithful cone forward projection (SW or HW clip box).  Returns (sino, bound|None, n_samples)."""
    vol = _c(vol, np.float32)
    D, H, W = vol.shape
    inv_ar = _c(inv_ar, np.float32).reshape(-1, 9)
    src = _c(src, np.float32).reshape(-1, 3)
    n = inv_ar.shape[0]
    v0, v1 = (0, n) if views is None else views
    Hd, Wd = det_shape
    out = np.empty((v1 - v0, Hd, Wd), np.float32)
    bound = np.empty_like(out) if want_bound else None
    ns = ctypes.c_longlong(0)
    lib().orc_cone_fwd_f32(vol, D, H, W, inv_ar, src, n, v0, v1, Hd, Wd, _c(spacing_zyx, np.float32),
                           float(step), int(bool(hw)), int(qbits), out, _ptr(bound), ctypes.byref(ns))
    return out, bound, ns.value


def cone_forward_f64(vol, inv_ar64, src64, det_shape, spacing_zyx, step, hw=False, views=None):
    vol = _c(vol, np.float64)
    D, H, W = vol.shape
    inv_ar64 = _c(inv_ar64, np.float64).reshape(-1, 9)
    src64 = _c(src64, np.float64).reshape(-1, 3)
    n = inv_ar64.shape[0]
    v0, v1 = (0, n) if views is None else views
    Hd, Wd = det_shape
    out = np.empty((v1 - v0, Hd, Wd), np.float64)
    ns = ctypes.c_longlong(0)
    lib().orc_cone_fwd_f64(vol, D, H, W, inv_ar64, src64, n, v0, v1, Hd, Wd, _c(spacing_zyx, np.float64),
                           float(step), int(bool(hw)), out, ctypes.byref(ns))
    return out, ns.value


def cone_backward(sino, P, vol_shape, spacing_zyx, origin_zyx, mult=1.0, hw=False, qbits=0, quirk=False,
                  zrange=None, want_bound=False):
    """f32-faithful voxel-driven cone backprojection.  Returns (vol[z1-z0,H,W], bound|None)."""
    sino = _c(sino, np.float32)
    n, Hd, Wd = sino.shape
    P = _c(P, np.float32).reshape(-1, 12)
    D, H, W = vol_shape
    z0, z1 = (0, D) if zrange is None else zrange
    out = np.empty((z1 - z0, H, W), np.float32)
    bound = np.empty_like(out) if want_bound else None
    lib().orc_cone_bp_f32(sino, n, Hd, Wd, P, D, H, W, _c(spacing_zyx, np.float32), _c(origin_zyx, np.float32),
                          float(mult), int(bool(hw)), int(qbits), int(bool(quirk)), z0, z1, out, _ptr(bound))
    return out, bound


def cone_backward_f64(sino, P, vol_shape, spacing_zyx, origin_zyx, mult=1.0, hw=False, zrange=None):
    sino = _c(sino, np.float64)
    n, Hd, Wd = sino.shape
    P = _c(P, np.float64).reshape(-1, 12)
    D, H, W = vol_shape
    z0, z1 = (0, D) if zrange is None else zrange
    out = np.empty((z1 - z0, H, W), np.float64)
    lib().orc_cone_bp_f64(sino, n, Hd, Wd, P, D, H, W, _c(spacing_zyx, np.float64), _c(origin_zyx, np.float64),
                          float(mult), int(bool(hw)), z0, z1, out)
    return out


# --------------------------------------------------------------------------- 2D


def _proj2d(fan, vol, rays, det, vol_origin_yx, vol_spacing_yx, det_origin, det_spacing, sid, sdd,
            step_factor, qbits, want_bound, f64):
    if f64:
        vol = _c(vol, np.float64)
        rays = _c(rays, np.float64).reshape(-1, 2)
        H, W = vol.shape
        n = rays.shape[0]
        out = np.empty((n, det), np.float64)
        ns = ctypes.c_longlong(0)
        lib().orc_proj2d_f64(int(fan), vol, H, W, rays, n, det, _c(vol_origin_yx, np.float64),
                             _c(vol_spacing_yx, np.float64), float(det_origin), float(det_spacing), float(sid),
                             float(sdd), float(step_factor), out, ctypes.byref(ns))
        return out, None, ns.value
    vol = _c(vol, np.float32)
    rays = _c(rays, np.float32).reshape(-1, 2)
    H, W = vol.shape
    n = rays.shape[0]
    out = np.empty((n, det), np.float32)
    bound = np.empty_like(out) if want_bound else None
    ns = ctypes.c_longlong(0)
    lib().orc_proj2d_f32(int(fan), vol, H, W, rays, n, det, _c(vol_origin_yx, np.float32),
                         _c(vol_spacing_yx, np.float32), float(det_origin), float(det_spacing), float(sid),
                         float(sdd), float(step_factor), int(qbits), out, _ptr(bound), ctypes.byref(ns))
    return out, bound, ns.value


def parallel_forward(vol, rays, det, vol_origin_yx, vol_spacing_yx, det_origin, det_spacing, step_factor=0.2,
                     qbits=0, want_bound=False, f64=False):
    return _proj2d(0, vol, rays, det, vol_origin_yx, vol_spacing_yx, det_origin, det_spacing, 0.0, 0.0,
                   step_factor, qbits, want_bound, f64)


def fan_forward(vol, rays, det, vol_origin_yx, vol_spacing_yx, det_origin, det_spacing, sid, sdd,
                step_factor=0.2, qbits=0, want_bound=False, f64=False):
    return _proj2d(1, vol, rays, det, vol_origin_yx, vol_spacing_yx, det_origin, det_spacing, sid, sdd,
                   step_factor, qbits, want_bound, f64)


def _back2d(fan, sino, rays, vol_shape, vol_origin_yx, vol_spacing_yx, det_origin, det_spacing, sid, sdd,
            qbits, want_bound, f64):
    H, W = vol_shape
    if f64:
        sino = _c(sino, np.float64)
        n, det = sino.shape
        out = np.empty((H, W), np.float64)
        lib().orc_back2d_f64(int(fan), sino, n, det, _c(rays, np.float64).reshape(-1, 2), H, W,
                             _c(vol_origin_yx, np.float64), _c(vol_spacing_yx, np.float64), float(det_origin),
                             float(det_spacing), float(sid), float(sdd), out)
        return out, None
    sino = _c(sino, np.float32)
    n, det = sino.shape
    out = np.empty((H, W), np.float32)
    bound = np.empty_like(out) if want_bound else None
    lib().orc_back2d_f32(int(fan), sino, n, det, _c(rays, np.float32).reshape(-1, 2), H, W,
                         _c(vol_origin_yx, np.float32), _c(vol_spacing_yx, np.float32), float(det_origin),
                         float(det_spacing), float(sid), float(sdd), int(qbits), out, _ptr(bound))
    return out, bound


def parallel_backward(sino, rays, vol_shape, vol_origin_yx, vol_spacing_yx, det_origin, det_spacing, qbits=0,
                      want_bound=False, f64=False):
    return _back2d(0, sino, rays, vol_shape, vol_origin_yx, vol_spacing_yx, det_origin, det_spacing, 0.0, 0.0,
                   qbits, want_bound, f64)


def fan_backward(sino, rays, vol_shape, vol_origin_yx, vol_spacing_yx, det_origin, det_spacing, sid, sdd,
                 qbits=0, want_bound=False, f64=False):
    return _back2d(1, sino, rays, vol_shape, vol_origin_yx, vol_spacing_yx, det_origin, det_spacing, sid, sdd,
                   qbits, want_bound, f64)
