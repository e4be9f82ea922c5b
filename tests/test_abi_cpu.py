"""C ABI checks that need no GPU: exported symbols, argument validation, op registration."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from pyronn_layers import native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(pyronn_[a-z0-9_]+)\s*\(", txt))
    return names


def test_every_declared_symbol_is_exported():
    declared = _declared()
    assert len(declared) >= 11
    L = nat.lib()
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert declared <= set(nat.EXPORTED)


def test_version_and_empty_error():
    assert nat.lib().pyronn_version() >= 100
    assert nat.last_error() == ""


def _fake(n=1):
    return ctypes.c_void_p(0x1000 * n)


def test_invalid_arguments_are_rejected_before_launch():
    L = nat.lib()
    # batch = 0
    rc = L.pyronn_cone_projection3d(_fake(), 0, 0, 4, 4, 4, _fake(), _fake(), _fake(), _fake(), _fake(), 2, 4, 4, 0,
                                    _fake(), _fake(), 1 << 20, None)
    assert rc == 1 and "bad shape" in nat.last_error()
    # bad dtype
    rc = L.pyronn_cone_projection3d(_fake(), 7, 1, 4, 4, 4, _fake(), _fake(), _fake(), _fake(), _fake(), 2, 4, 4, 0,
                                    _fake(), _fake(), 1 << 20, None)
    assert rc == 1 and "dtype" in nat.last_error()
    # null pointer
    rc = L.pyronn_cone_backprojection3d(None, 0, 1, 2, 4, 4, _fake(), _fake(), _fake(), _fake(), _fake(), 4, 4, 4,
                                        0, _fake(), _fake(), 1 << 20, None)
    assert rc == 1 and "null" in nat.last_error()
    # workspace too small
    need = L.pyronn_cone_projection3d_workspace_size(1, 4, 4, 4, 2, 0)
    assert need > 0
    rc = L.pyronn_cone_projection3d(_fake(), 0, 1, 4, 4, 4, _fake(), _fake(), _fake(), _fake(), _fake(), 2, 4, 4, 0,
                                    _fake(), _fake(), need - 1, None)
    assert rc == 2 and "workspace" in nat.last_error()
    # 2D: non-positive step factor
    rc = L.pyronn_parallel_projection2d(_fake(), 0, 1, 8, 8, _fake(), _fake(), _fake(), _fake(), _fake(), 4, 9, 0.0,
                                        _fake(), None)
    assert rc == 1 and "step" in nat.last_error()
    rc = L.pyronn_fan_projection2d(_fake(), 0, 1, 8, 8, _fake(), _fake(), _fake(), _fake(), None, _fake(), _fake(), 4,
                                   9, 0.2, _fake(), None)
    assert rc == 1 and "source" in nat.last_error()


def test_workspace_alignment_is_checked():
    L = nat.lib()
    need = L.pyronn_cone_projection3d_workspace_size(1, 4, 4, 4, 2, 0)
    rc = L.pyronn_cone_projection3d(_fake(), 0, 1, 4, 4, 4, _fake(), _fake(), _fake(), _fake(), _fake(), 2, 4, 4, 0,
                                    _fake(), ctypes.c_void_p(0x1008), need, None)
    assert rc == 1 and "aligned" in nat.last_error()
    need = L.pyronn_cone_backprojection3d_workspace_size(1, 2, 4, 4, 0)
    rc = L.pyronn_cone_backprojection3d(_fake(), 0, 1, 2, 4, 4, _fake(), _fake(), _fake(), _fake(), _fake(), 4, 4, 4,
                                        0, _fake(), ctypes.c_void_p(0x1004), need, None)
    assert rc == 1 and "aligned" in nat.last_error()


def test_view_range_backprojection_validation():
    L = nat.lib()
    args = lambda mode, ws, z0=0, z1=4: (_fake(), 0, 1, 4, 8, 8, _fake(), _fake(), _fake(), _fake(), 4, 4, 4, 0, mode,
                                         z0, z1, _fake(), _fake(), ws, None)
    assert L.pyronn_cone_backprojection3d_views(*args(8, 1 << 20)) == 1 and "mode" in nat.last_error()
    assert L.pyronn_cone_backprojection3d_views(*args(-1, 1 << 20)) == 1
    for z0, z1 in ((-1, 4), (2, 2), (3, 1), (0, 5)):
        assert L.pyronn_cone_backprojection3d_views(*args(0, 1 << 20, z0, z1)) == 1
        assert "z range" in nat.last_error()
    need = L.pyronn_cone_backprojection3d_workspace_size(1, 4, 8, 8, 0)
    assert L.pyronn_cone_backprojection3d_views(*args(3, need - 1)) == 2
    rc = L.pyronn_cone_backprojection3d_views(None, 0, 1, 4, 8, 8, _fake(), _fake(), _fake(), _fake(), 4, 4, 4, 0, 0,
                                              0, 4, _fake(), _fake(), need, None)
    assert rc == 1 and "null" in nat.last_error()


def _fwd_rc(L, D, H, W, n=2, det=(4, 4), ws=16, batch=1):
    return L.pyronn_cone_projection3d(_fake(), 0, batch, D, H, W, _fake(), _fake(), _fake(), _fake(), _fake(), n,
                                      det[0], det[1], 0, _fake(), _fake(), ws, None)


def _bp_rc(L, D, H, W, n=2, det=(4, 4), ws=16, batch=1):
    return L.pyronn_cone_backprojection3d(_fake(), 0, batch, n, det[0], det[1], _fake(), _fake(), _fake(), _fake(),
                                          _fake(), D, H, W, 0, _fake(), _fake(), ws, None)


# Volumes of fewer than 2^32 voxels (the reference's uint32 index limit,
# cc/kernels/cone_projector_3D_CudaKernel.cu.cc:52-59): cubes, slabs, rods, planes and
# degenerate shapes, each with every side below 2^24 - 3.
ACCEPTED_SHAPES = [(1625, 1625, 1625), (64, 32768, 8), (8, 32768, 64), (2048, 2048, 1023), (1, 1, (1 << 24) - 4),
                   ((1 << 24) - 4, 1, 1), (1, (1 << 24) - 4, 1), (1, 65535, 65535), (65535, 65535, 1),
                   (256, 4096, 4095), (1287, 1287, 1287), (1300, 1300, 1300)]


@pytest.mark.parametrize("shape", ACCEPTED_SHAPES)
def test_volumes_below_2_32_voxels_are_accepted(shape):
    """Every check up to the workspace size passes (status 2 = workspace too small, with a
    16-byte workspace): no PYRONN_ERR_TOO_LARGE for any volume the reference accepts."""
    D, H, W = shape
    assert D * H * W < 2 ** 32
    L = nat.lib()
    assert _fwd_rc(L, D, H, W) == 2, nat.last_error()
    assert "workspace" in nat.last_error()
    assert _bp_rc(L, D, H, W) == 2, nat.last_error()
    assert "workspace" in nat.last_error()
    # the workspace sizes are computed in 64-bit: (W+2)(H+2)(D+3) xy quads for the forward
    need = L.pyronn_cone_projection3d_workspace_size(1, D, H, W, 2, 0)
    assert need >= (W + 2) * (H + 2) * (D + 3) * 16


def test_views_and_batches_beyond_one_grid_are_accepted():
    """A 2048x2048 detector needs 16384 workgroups per view; 2000 views (3.3e7 workgroups)
    exceed one dispatch's 2^32 work-items and run as several launches."""
    L = nat.lib()
    assert _fwd_rc(L, 64, 64, 64, n=2000, det=(2048, 2048)) == 2, nat.last_error()
    assert _bp_rc(L, 64, 64, 64, n=2000, det=(2048, 2048)) == 2, nat.last_error()
    # batch x depth beyond 65535 z blocks: split into launches
    assert _bp_rc(L, 4096, 16, 16, batch=300) == 2, nat.last_error()


def test_too_large_volume_rejected():
    L = nat.lib()
    # a side of 2^24 - 3 voxels or more (coordinates would no longer be exact floats)
    assert _fwd_rc(L, 1, 1, (1 << 24) - 3, ws=1 << 62) == 3
    assert "2^24" in nat.last_error()
    assert _fwd_rc(L, (1 << 24) + 5, 1, 1, ws=1 << 62) == 3
    # a detector image beyond the backprojection's 31-bit cell offsets
    assert _bp_rc(L, 4, 4, 4, det=(1 << 15, 1 << 15), ws=1 << 62) == 3
    assert "detector" in nat.last_error()
    # batch x views beyond int
    assert _fwd_rc(L, 4, 4, 4, n=1 << 20, batch=4096, ws=1 << 62) == 1


def test_workspace_sizes_scale_with_storage():
    L = nat.lib()
    f32 = L.pyronn_cone_projection3d_workspace_size(1, 64, 64, 64, 10, 0)
    bf16 = L.pyronn_cone_projection3d_workspace_size(1, 64, 64, 64, 10, 1)
    assert f32 > bf16 > 0
    # per view: (H_d + 2) x (W_d + 2) cells of 4 neighbours
    assert L.pyronn_cone_backprojection3d_workspace_size(2, 10, 8, 8, 0) == \
        ((2 * 10 * 10 * 10 * 4 * 4 + 255) // 256) * 256
    # geometry (12 floats per view) + cells ((W+2)(H+2)(D+3) xy quads); no other copy of the volume
    assert L.pyronn_cone_projection3d_workspace_size(1, 4, 4, 4, 2, 1) == \
        ((1 * 2 * 12 * 4 + 255) // 256) * 256 + ((6 * 6 * 7 * 4 * 2 + 255) // 256) * 256
    # the 16-bit copy is half the bytes of the f32 one
    assert L.pyronn_cone_projection3d_workspace_size(1, 4, 4, 4, 2, 0) == \
        ((1 * 2 * 12 * 4 + 255) // 256) * 256 + ((6 * 6 * 7 * 4 * 4 + 255) // 256) * 256


def test_torch_ops_registered_and_cpu_tensors_rejected():
    import pyronn_layers as pl
    pl.load()
    for name in ["parallel_projection2d", "parallel_backprojection2d", "fan_projection2d",
                 "fan_backprojection2d", "cone_projection3d", "cone_backprojection3d"]:
        assert hasattr(torch.ops.pyronn_layers, name)
    vol = torch.zeros(1, 4, 4, 4)
    with pytest.raises(Exception):
        pl.cone_projection3d(vol, [2, 4, 4], torch.zeros(1, 3), torch.ones(1, 3), torch.zeros(1, 2, 3, 4),
                             torch.ones(1, 1), torch.ones(1, 1))


def test_batched_argument_validation():
    L = nat.lib()
    P = nat.ConeProblem
    good = P(0x1000, 0x2000, 0x3000, 0x4000, 0x5000, 0x6000, 0x7000, 4, 4, 4, 2, 4, 4)
    arr = (P * 17)(*([good] * 17))
    assert L.pyronn_cone_projection3d_batched(arr, 0, 0, 0, 0, 0, _fake(), 1 << 30, None) == 1
    assert "count" in nat.last_error()
    assert L.pyronn_cone_projection3d_batched(arr, 17, 0, 0, 0, 0, _fake(), 1 << 30, None) == 1
    bad = (P * 2)(good, P(0, 0x2000, 0x3000, 0x4000, 0x5000, 0x6000, 0x7000, 4, 4, 4, 2, 4, 4))
    assert L.pyronn_cone_projection3d_batched(bad, 2, 0, 0, 0, 0, _fake(), 1 << 30, None) == 1
    assert "problem 1" in nat.last_error()
    assert L.pyronn_cone_backprojection3d_batched(arr, 2, 0, 0, 5, 0, _fake(), 1 << 30, None) == 1
    assert "z_per_lane" in nat.last_error()
    need = L.pyronn_cone_projection3d_batched_workspace_size(arr, 2, 0)
    assert need == 256 + 2 * L.pyronn_cone_projection3d_workspace_size(1, 4, 4, 4, 2, 0)
    assert L.pyronn_cone_projection3d_batched(arr, 2, 0, 0, 0, 0, _fake(), need - 1, None) == 2
    # forward_tile outside 0..3
    assert L.pyronn_cone_projection3d_batched(arr, 2, 0, 0, 0, 4, _fake(), need, None) == 1
    assert "forward_tile" in nat.last_error()
    assert L.pyronn_cone_projection3d_batched(arr, 2, 0, 0, 0, -1, _fake(), need, None) == 1


def test_fft_filter_limits_and_validation():
    L = nat.lib()
    # table floats: ramp spectrum (L) + 4 per twiddle pair + column weights (W), L = pow2 >=
    # max(16, 2 W - 1); one twiddle pair per butterfly index k < Ns of each pass with Ns > 1,
    # forward and inverse (radix 2 or 4 first when log2 L is not a multiple of 3, radix 8 after):
    #   L = 16   (2, 8):        forward Ns 2; inverse (8, 2) Ns 8                     -> 10 pairs
    #   L = 32   (4, 8):        forward Ns 4; inverse Ns 8                            -> 12
    #   L = 2048 (4, 8, 8, 8):  forward Ns 4, 32, 256; inverse Ns 8, 64, 512          -> 876
    #   L = 8192 (2, 8, 8, 8, 8): forward 2, 16, 128, 1024; inverse 8, 64, 512, 4096  -> 5850
    assert L.pyronn_fdk_ramp_spectrum_size(1) == 16 + 4 * 10 + 1
    assert L.pyronn_fdk_ramp_spectrum_size(9) == 32 + 4 * 12 + 9
    assert L.pyronn_fdk_ramp_spectrum_size(1024) == 2048 + 4 * 876 + 1024
    assert L.pyronn_fdk_ramp_spectrum_size(4096) == 8192 + 4 * 5850 + 4096
    assert L.pyronn_fdk_ramp_spectrum_size(4097) == 0
    assert L.pyronn_fdk_ramp_spectrum_size(0) == 0
    assert L.pyronn_fdk_ramp_spectrum(4097, 1.0, None, _fake(), None) == 1 and "outside" in nat.last_error()
    assert L.pyronn_fdk_ramp_spectrum(64, 0.0, None, _fake(), None) == 1 and "spacing" in nat.last_error()
    for addr in (0x1004, 0x1008):   # the table is read with 16-byte loads
        assert L.pyronn_fdk_ramp_spectrum(64, 1.0, None, ctypes.c_void_p(addr), None) == 1
        assert "16-byte aligned" in nat.last_error()
    args = lambda imgs, h, w, dt: (_fake(), dt, imgs, h, w, None, _fake(2), _fake(3), 0, None)
    assert L.pyronn_fdk_filter_fft(*args(1, 4, 5000, 0)) == 1 and "outside" in nat.last_error()
    assert L.pyronn_fdk_filter_fft(*args(0, 4, 64, 0)) == 1 and "shape" in nat.last_error()
    assert L.pyronn_fdk_filter_fft(*args(1, 4, 64, 9)) == 1 and "dtype" in nat.last_error()
    assert L.pyronn_fdk_filter_fft(*args(1 << 30, 8, 64, 0)) == 1 and "pairs" in nat.last_error()
    assert L.pyronn_fdk_filter_fft(None, 0, 1, 4, 64, None, _fake(2), _fake(3), 0, None) == 1
    assert "null" in nat.last_error()



def test_noalloc_forward_checks_arguments_before_any_device_call():
    """The no-allocation forward rejects bad arguments and more views than its fixed scratch
    holds (batch x views > 65536) on the host, before touching the device."""
    L = nat.lib()
    rc = L.pyronn_cone_projection3d_noalloc(_fake(), 0, 2, 4, 4, 4, _fake(), _fake(), _fake(), _fake(), _fake(),
                                            40000, 4, 4, 0, _fake(), None)
    assert rc == 3 and "scratch" in nat.last_error()
    rc = L.pyronn_cone_projection3d_noalloc(None, 0, 1, 4, 4, 4, _fake(), _fake(), _fake(), _fake(), _fake(), 2, 4,
                                            4, 0, _fake(), None)
    assert rc == 1 and "null" in nat.last_error()
    rc = L.pyronn_cone_projection3d_noalloc(_fake(), 5, 1, 4, 4, 4, _fake(), _fake(), _fake(), _fake(), _fake(), 2,
                                            4, 4, 0, _fake(), None)
    assert rc == 1 and "dtype" in nat.last_error()
