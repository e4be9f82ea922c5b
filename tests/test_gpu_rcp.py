"""Exhaustive device checks of two arithmetic shortcuts the kernels rely on.

* The backprojector divides by the projected depth with rcp_rn() (hardware reciprocal
  estimate + two FMA Newton corrections, IEEE division outside [2^-125, 2^125]).  It is
  checked against the IEEE division 1.0f / w for all 2^32 float bit patterns, so the
  backprojection keeps the reference's correctly rounded division.
* The samplers clamp coordinates into the zero border with one v_med3_f32.  It is
  checked against the oracle's fminf(fmaxf(c, -1), n) for all 2^32 inputs except
  signaling NaNs (which no arithmetic produces), for several border sizes n."""
import ctypes
import os

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

KIT = os.path.join(os.path.dirname(__file__), "native", "libpyronn_testkit.so")


def _kit():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(KIT):
        pytest.fail("tests/native/libpyronn_testkit.so missing: run pyro-nn-layers_amd/build.py")
    kit = ctypes.CDLL(KIT)
    kit.pyronn_testkit_rcp_check.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    kit.pyronn_testkit_rcp_check.restype = ctypes.c_int
    kit.pyronn_testkit_clamp_check.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p]
    kit.pyronn_testkit_clamp_check.restype = ctypes.c_int
    return kit


def test_rcp_rn_matches_ieee_division_for_every_float():
    kit = _kit()
    counts = torch.full((1 << 24,), 0xFFFF, dtype=torch.int32, device="cuda")
    rc = kit.pyronn_testkit_rcp_check(ctypes.c_void_p(counts.data_ptr()),
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    bad = int(counts.sum().item())
    print(f"rcp_rn vs 1.0f/w: {bad} mismatches over 2^32 inputs")
    assert bad == 0


@pytest.mark.parametrize("n", [1.0, 256.0, 767.0, 1024.0, 8388608.0])
def test_med3_clamp_matches_min_max_for_every_float(n):
    kit = _kit()
    counts = torch.full((1 << 24,), 0xFFFF, dtype=torch.int32, device="cuda")
    rc = kit.pyronn_testkit_clamp_check(ctypes.c_void_p(counts.data_ptr()), n,
                                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    bad = int(counts.sum().item())
    print(f"clamp_apron(c, {n}) vs fminf(fmaxf(c, -1), n): {bad} mismatches over 2^32 - 2^22 inputs")
    assert bad == 0
