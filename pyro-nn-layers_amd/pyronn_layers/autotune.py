"""Project-written autotuner for the batched persistent launches.

Tunable parameters (all preserve results bit for bit):
  forward          forward_tile in {0, 1, 2, 3}: the detector tile of a workgroup and its
                   waves (16x16 of 4x16 waves, 8x32 of 2x32, 32x8 of 8x8, 16x16 of 8x8),
                   timed with one workgroup per item; then blocks_per_cu in {8, 16} for the
                   best tile (a persistent grid of CUs x value workgroups)
  backprojection   z_per_lane in {4, 8} x blocks_per_cu in {0, 4, 8, 16}
                   (0: one workgroup per item; otherwise a persistent grid)
Each candidate is timed with HIP events (best of `reps` after one warm-up) and
the winner is cached in a JSON file keyed by the batch signature (shapes,
views, detector sizes, storage dtype, interpolation mode, device name), so a later
call with the same signature reuses it without timing.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Sequence

import torch

from . import batched
from . import geometry as geo

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "pyronn_mi_autotune.json")


TILE_SHAPES = {0: "16x16/4x16", 1: "8x32/2x32", 2: "32x8/8x8", 3: "16x16/8x8"}   # tile / wave (cols x rows)


def signature(kind: str, tensors: Sequence[torch.Tensor], geometries: Sequence[geo.ConeGeometry],
              hardware_interp: bool = False) -> str:
    dev = torch.cuda.get_device_name(tensors[0].device)
    parts = [f"{tuple(g.volume_shape)}:{g.n_views}:{tuple(g.detector_shape)}" for g in geometries]
    return f"{kind}|{tensors[0].dtype}|hw={int(bool(hardware_interp))}|{dev}|" + ";".join(parts)


def _time(fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


class Autotuner:
    def __init__(self, cache_path: str = DEFAULT_CACHE, reps: int = 2):
        self.path = cache_path
        self.reps = reps
        self.cache: Dict[str, dict] = {}
        if cache_path and os.path.exists(cache_path):
            try:
                with open(cache_path) as f:
                    self.cache = json.load(f)
            except (OSError, ValueError):
                self.cache = {}

    def _save(self):
        if not self.path:
            return
        os.makedirs(os.path.dirname(self.path), exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.cache, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def forward(self, volumes, geometries, hardware_interp=False) -> dict:
        key = signature("fwd", volumes, geometries, hardware_interp)
        if key not in self.cache:
            trials = {}

            def trial(bpc, tile):
                trials[(bpc, tile)] = _time(lambda: batched.cone_projection3d_batched(
                    volumes, geometries, hardware_interp, blocks_per_cu=bpc, forward_tile=tile), self.reps)

            for tile in TILE_SHAPES:
                trial(0, tile)
            tile = min(trials, key=trials.get)[1]
            for bpc in (8, 16):
                trial(bpc, tile)
            best = min(trials, key=trials.get)
            self.cache[key] = {"blocks_per_cu": best[0], "forward_tile": best[1], "tile_shape": TILE_SHAPES[best[1]],
                               "ms": trials[best],
                               "trials_ms": {f"{k[0]}/{TILE_SHAPES[k[1]]}": v for k, v in trials.items()}}
            self._save()
        return self.cache[key]

    def backward(self, sinograms, geometries, hardware_interp=False) -> dict:
        key = signature("bp", sinograms, geometries, hardware_interp)
        if key not in self.cache:
            trials = {}
            for zpl in (4, 8):
                for bpc in (0, 4, 8, 16):
                    trials[(zpl, bpc)] = _time(lambda: batched.cone_backprojection3d_batched(
                        sinograms, geometries, hardware_interp, z_per_lane=zpl, blocks_per_cu=bpc), self.reps)
            best = min(trials, key=trials.get)
            self.cache[key] = {"z_per_lane": best[0], "blocks_per_cu": best[1], "ms": trials[best],
                               "trials_ms": {f"{k[0]}x{k[1]}": v for k, v in trials.items()}}
            self._save()
        return self.cache[key]
