"""The closed-form forward sample count used by bench.py vs the oracle's exact count."""
import numpy as np
import pytest

import oracle as orc
from pyronn_layers import geometry as geo
from pyronn_layers import metrics


@pytest.mark.parametrize("hw", [False, True])
def test_sample_count_matches_oracle(hw):
    shape = (20, 24, 28)
    sp = np.array([1.1, 1.0, 0.9], np.float32)
    og = geo.centred_origin(shape, sp).astype(np.float32)
    P = geo.circular_projection_matrices(6, 80.0, 140.0, (18, 22), (1.5, 1.5), start=0.2).astype(np.float32)
    inv, src, _, _, _ = orc.cone_setup(P, sp, og)
    _, _, exact = orc.cone_forward(np.zeros(shape, np.float32), inv, src, (18, 22), sp, 0.5, hw=hw)
    approx = metrics.cone_forward_sample_count(P, shape, sp, og, (18, 22), 0.5, hardware_interp=hw, device="cpu")
    assert exact > 10000
    assert abs(approx - exact) <= max(2, 1e-5 * exact)


def test_bench_refuses_multi_gpu():
    import bench
    assert bench.main(["--gpus", "2"]) == 2
