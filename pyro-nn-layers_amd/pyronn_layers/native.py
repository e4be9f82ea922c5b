"""ctypes binding of the C ABI in include/pyronn_mi.h (libpyronn_mi.so).

This is the framework-neutral way into the kernels: every argument is a raw
device pointer (int), a host integer or a stream handle.  The torch op library
is the other way in; the parity tests use this one.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(_HERE, "lib", "libpyronn_mi.so")

PYRONN_OK = 0
F32, BF16, F16 = 0, 1, 2
MAX_BATCHED_PROBLEMS = 16
FORWARD_TILES = 4   # forward_tile values of pyronn_cone_projection3d_batched


class ConeProblem(ctypes.Structure):
    """struct pyronn_cone_problem (include/pyronn_mi.h)."""
    _fields_ = [("input", ctypes.c_void_p), ("output", ctypes.c_void_p), ("volume_origin", ctypes.c_void_p),
                ("volume_spacing", ctypes.c_void_p), ("projection_matrices", ctypes.c_void_p),
                ("step_size", ctypes.c_void_p), ("projection_multiplier", ctypes.c_void_p),
                ("D", ctypes.c_int), ("H", ctypes.c_int), ("W", ctypes.c_int), ("n_views", ctypes.c_int),
                ("det_h", ctypes.c_int), ("det_w", ctypes.c_int)]


_lib = None
_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_sz = ctypes.c_size_t

_SIGS = {
    "pyronn_last_error": (ctypes.c_char_p, []),
    "pyronn_version": (_i, []),
    "pyronn_cone_geometry_setup": (_i, [_vp, _vp, _vp, _i, _i, _vp, _vp, _vp]),
    "pyronn_cone_projection3d_workspace_size": (_sz, [_i, _i, _i, _i, _i, _i]),
    "pyronn_cone_projection3d": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp,
                                      _vp, _sz, _vp]),
    "pyronn_cone_projection3d_noalloc": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i,
                                              _vp, _vp]),
    "pyronn_cone_backprojection3d_workspace_size": (_sz, [_i, _i, _i, _i, _i]),
    "pyronn_cone_backprojection3d": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i,
                                          _vp, _vp, _sz, _vp]),
    "pyronn_cone_backprojection3d_views": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _i, _i,
                                                _i, _i, _i, _vp, _vp, _sz, _vp]),
    "pyronn_parallel_projection2d": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _f, _vp,
                                          _vp]),
    "pyronn_parallel_backprojection2d": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp,
                                              _vp]),
    "pyronn_fan_projection2d": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _f,
                                     _vp, _vp]),
    "pyronn_fan_backprojection2d": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i,
                                         _vp, _vp]),
    "pyronn_cone_projection3d_batched_workspace_size": (_sz, [_vp, _i, _i]),
    "pyronn_cone_projection3d_batched": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _sz, _vp]),
    "pyronn_cone_backprojection3d_batched_workspace_size": (_sz, [_vp, _i, _i]),
    "pyronn_cone_backprojection3d_batched": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _sz, _vp]),
    "pyronn_fdk_ramp_matrix": (_i, [_i, ctypes.c_double, _vp, _vp, _vp]),
    "pyronn_fdk_cosine_weights": (_i, [_i, _i] + [ctypes.c_double] * 5 + [_vp, _vp]),
    "pyronn_fdk_filter_workspace_size": (_sz, [_i, _i, _i]),
    "pyronn_fdk_filter": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _sz, _vp]),
    "pyronn_fdk_ramp_spectrum_size": (_sz, [_i]),
    "pyronn_fdk_ramp_spectrum": (_i, [_i, ctypes.c_double, _vp, _vp, _vp]),
    "pyronn_fdk_filter_fft": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp]),
}

EXPORTED = tuple(_SIGS)


class PyronnError(RuntimeError):
    pass


def lib(path: str = HIP_LIB):
    """Load libpyronn_mi.so (raises if it is missing; there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise PyronnError(f"{path} not built; run `python pyro-nn-layers_amd/build.py`")
        L = ctypes.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().pyronn_last_error().decode()


def check(rc: int, what: str = "") -> None:
    if rc != PYRONN_OK:
        raise PyronnError(f"{what} failed (status {rc}): {last_error()}")


def call(name: str, *args) -> int:
    """Call an entry point; raises PyronnError on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    check(rc, name)
    return rc
