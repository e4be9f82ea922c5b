"""CPU checks of two integer facts the 16-bit cone forward relies on (pyro-nn-layers_amd/csrc/cone.hip).

* The wave-uniform sample count of the streamed march (march_ray_stream): from the alpha
  interval [a0, amax) of a ray, safe = int((amax - a0) / (step + (M + step) 2^-22)) - 2 with
  M = max(|a0|, |amax|), all in float32.  The march then takes that many alpha += step
  increments without testing alpha < amax, so safe must never exceed the number of increments
  the tested loop of cone_projector_3D_CudaKernel.cu.cc:134-162 takes.  Checked here against that
  loop run in float32 (brute force) over random intervals and over the adversarial corner of
  the unguarded march's domain (|alpha| near 2^19, step near 2^-20 |alpha|).
* The byte offset of a 16-bit cell (locate16, IT = unsigned): m = fma(fma(fy, Dc, fz), 8,
  2^23 + 8 (Dc + 1)) is an exact integer in [2^23, 2^24), so its float32 bit pattern is
  0x4B000000 + m - 2^23; the 24-bit multiply-add fx * 8 HDc + bits, minus the bias
  0x4B000000 - 8 HDc, is 8 x the cell index ((fx+1)(H+2) + (fy+1)) Dc + fz + 1, and stays in
  [0, 2^32) for every floor in [-1, N] of every shape fwd_byte_offsets admits.  Checked at the
  corners of the coordinate box for shapes at and beyond the limits.
"""
import itertools

import numpy as np
import pytest

f32 = np.float32


def _safe(a0, amax, step):
    mag = max(abs(a0), abs(amax))
    adv = f32(step + f32(f32(mag + step) * f32(2.0 ** -22)))
    return int(f32(f32(amax - a0) / adv)) - 2


def _taken(a0, amax, step):
    a, n = a0, 0
    while True:
        a = f32(a + step)
        if not a < amax:
            return n
        n += 1


def _check(a0, amax, step):
    a0, amax, step = f32(a0), f32(amax), f32(step)
    if not a0 < amax:
        return 0
    s, n = _safe(a0, amax, step), _taken(a0, amax, step)
    assert s <= n, (float(a0), float(amax), float(step), s, n)
    return n - s


def test_uniform_count_never_exceeds_the_tested_march():
    rng = np.random.default_rng(7)
    slack = []
    for _ in range(1500):
        mag = 2.0 ** rng.uniform(-3, 19)
        step = mag * 2.0 ** rng.uniform(-20, -5)
        a0 = rng.uniform(-mag, mag)
        slack.append(_check(a0, a0 + rng.uniform(0, 1) * min(2 * mag, step * 2000), step))
    # the count is tight where it matters: half-voxel steps through a volume of a few hundred voxels
    for _ in range(200):
        a0 = rng.uniform(500.0, 1500.0)
        slack.append(_check(a0, a0 + rng.uniform(100.0, 800.0), 0.5))
    assert np.median(slack[-200:]) <= 4


def test_uniform_count_at_the_edge_of_the_unguarded_domain():
    rng = np.random.default_rng(11)
    for _ in range(200):
        mag = 2.0 ** rng.uniform(10, 19) * 0.999
        step = float(f32(mag * 2.0 ** -20)) * rng.uniform(1.0, 1.5)   # the smallest step plain rays allow
        a0 = float(f32(mag * rng.choice([-1.0, 1.0]) * rng.uniform(0.5, 1.0)))
        _check(a0, a0 + step * rng.uniform(1, 3000), step)


def _byte_offsets_ok(D, H, W):
    """fwd_byte_offsets (cone.hip)."""
    plane = (H + 2) * (D + 3)
    return W + 2 < 2 ** 23 and 8 * plane < 2 ** 23 and 8 * (W + 2) * plane <= 0xB5000000


@pytest.mark.parametrize("shape", [(512, 512, 512), (1000, 1020, 368), (32, 32, 32), (1, 1, 1), (3, 1000, 1),
                                   (1021, 1021, 2), (257, 3, 4000)])
def test_byte_offset_is_eight_cell_indices(shape):
    D, H, W = shape
    if not _byte_offsets_ok(D, H, W):
        pytest.skip("takes the element index")
    Dc, HDc = D + 3, (H + 2) * (D + 3)
    cells = (W + 2) * HDc
    mag = f32(2.0 ** 23) + f32(8 * (Dc + 1))
    assert float(mag) == 2 ** 23 + 8 * (Dc + 1)
    for fx, fy, fz in itertools.product((-1, 0, 1, W - 1, W), (-1, 0, 1, H - 1, H), (-1, 0, 1, D - 1, D)):
        yz = f32(f32(fy) * f32(Dc) + f32(fz))            # one FMA, exact: a small integer
        m = f32(float(yz) * 8.0 + float(mag))             # the second FMA, exact
        assert float(m) == float(yz) * 8.0 + float(mag) and 2 ** 23 <= float(m) < 2 ** 24
        bits = int(np.array([m], np.float32).view(np.uint32)[0])
        true_off = fx * 8 * HDc + bits                      # the 24-bit multiply-add, before wrapping
        assert 0 <= true_off < 2 ** 32
        cell = (fx + 1) * HDc + (fy + 1) * Dc + (fz + 1)
        assert true_off - 0x4B000000 + 8 * HDc == 8 * cell
        assert 0 <= cell and cell + 1 < cells               # the 16-byte load (cells z, z + 1) stays inside


def test_byte_offsets_limit_is_the_last_admitted_array():
    # the largest arrays fwd_byte_offsets admits still keep offset + bias below 2^32
    D, H = 1000, 1020
    plane = (H + 2) * (D + 3)
    W = 0xB5000000 // (8 * plane) - 2
    assert _byte_offsets_ok(D, H, W) and not _byte_offsets_ok(D, H, W + 1)
    assert 8 * (W + 2) * plane + 0x4B000000 - 8 * plane <= 2 ** 32
    # and the first plane too wide for the 8-byte scale takes the element index
    assert not _byte_offsets_ok(1021, 1022, 2)
