"""The backprojectors divide by the projected depth with rcp_rn() (hardware reciprocal
estimate + two FMA Newton corrections, IEEE division outside [2^-125, 2^125]).  This
checks it against the IEEE division 1.0f / w on the device for all 2^32 float bit
patterns: the fast path must be bit-identical everywhere, so the backprojection keeps
the reference's correctly rounded division."""
import ctypes
import os

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

KIT = os.path.join(os.path.dirname(__file__), "native", "libpyronn_testkit.so")


def test_rcp_rn_matches_ieee_division_for_every_float():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(KIT):
        pytest.fail("tests/native/libpyronn_testkit.so missing: run pyro-nn-layers_amd/build.py")
    kit = ctypes.CDLL(KIT)
    kit.pyronn_testkit_rcp_check.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    kit.pyronn_testkit_rcp_check.restype = ctypes.c_int
    counts = torch.full((1 << 24,), 0xFFFF, dtype=torch.int32, device="cuda")
    rc = kit.pyronn_testkit_rcp_check(ctypes.c_void_p(counts.data_ptr()),
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    bad = int(counts.sum().item())
    print(f"rcp_rn vs 1.0f/w: {bad} mismatches over 2^32 inputs")
    assert bad == 0
