"""Continuous-batching reconstruction server (FDK / FBP requests of mixed sizes).

Requests arrive over time (e.g. a seeded Poisson process).  A request is
admitted as soon as (a) it has arrived, (b) an execution slot is free and
(c) its device-memory footprint fits in the HBM budget next to the requests
already running.  Each slot owns a HIP stream (default priority) and its own
device buffers, so admitted requests run concurrently on the device: the
upload of one overlaps the backprojection of another.  Plans (projection
matrices, cosine weights, ramp filter tables) are built once per geometry class and
stay resident; with `use_graphs` each slot replays a HIP graph captured per
class instead of launching the pipeline kernel by kernel.

The kernels are deterministic (no atomics; every output element is written by
exactly one lane with a fixed summation order), so a request's result does not
depend on what else is running: the same bits alone and in the mix.

Latency is measured from a request's arrival time to the moment its result is
back in host memory.
"""
from __future__ import annotations

import heapq
import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Any, Callable, Dict, List, Optional

import numpy as np


@dataclass
class Request:
    rid: int
    kind: str                 # geometry class key
    arrival: float            # seconds after the start of the run
    payload: Any = None       # executor-specific input (host sinogram)


@dataclass
class Result:
    rid: int
    kind: str
    arrival: float
    admitted: float = math.nan
    done: float = math.nan
    slot: int = -1
    output: Any = None
    error: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def poisson_arrivals(seed: int, rate_hz: float, count: int, kinds: Dict[str, float]) -> List[Request]:
    """`count` requests with exponential inter-arrival times (rate `rate_hz`) and
    classes drawn with the given probabilities; fully determined by `seed`."""
    rng = np.random.default_rng(seed)
    names = list(kinds)
    p = np.array([kinds[k] for k in names], np.float64)
    p /= p.sum()
    t = np.cumsum(rng.exponential(1.0 / rate_hz, size=count))
    cls = rng.choice(len(names), size=count, p=p)
    return [Request(i, names[c], float(t[i])) for i, c in enumerate(cls)]


class Scheduler:
    """Admission and completion bookkeeping; device work is delegated to an executor with
        footprint(kind) -> bytes, start(slot, request) -> handle,
        poll(handle) -> bool, finish(handle) -> output."""

    def __init__(self, executor, slots: int, hbm_budget: int):
        if slots < 1:
            raise ValueError("need at least one slot")
        self.ex = executor
        self.slots = slots
        self.budget = int(hbm_budget)

    def run(self, requests: List[Request], clock: Callable[[], float] = time.perf_counter,
            idle: Callable[[float], None] = None, realtime: bool = True) -> List[Result]:
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        free = list(range(self.slots))
        heapq.heapify(free)
        active = {}            # slot -> (request, handle, result, footprint)
        used = 0
        results: Dict[int, Result] = {}
        t0 = clock()
        while pending or active:
            now = clock() - t0
            # retire finished work, in slot order
            for slot in sorted(active):
                req, h, res, fp = active[slot]
                if self.ex.poll(h):
                    res.output = self.ex.finish(h)
                    res.done = clock() - t0
                    used -= fp
                    del active[slot]
                    heapq.heappush(free, slot)
            now = clock() - t0
            # admit in arrival order; a request that does not fit waits (FIFO, no overtaking)
            while pending and (pending[0].arrival <= now or not realtime) and free:
                req = pending[0]
                fp = self.ex.footprint(req.kind)
                if fp > self.budget:
                    pending.popleft()
                    results[req.rid] = Result(req.rid, req.kind, req.arrival, error=(
                        f"request needs {fp} bytes, more than the HBM budget {self.budget}"))
                    continue
                if used + fp > self.budget:
                    break
                pending.popleft()
                slot = heapq.heappop(free)
                res = Result(req.rid, req.kind, req.arrival, admitted=clock() - t0, slot=slot)
                try:
                    h = self.ex.start(slot, req)
                except Exception as e:  # executor refused: report, free the slot
                    res.error = str(e)
                    results[req.rid] = res
                    heapq.heappush(free, slot)
                    continue
                results[req.rid] = res
                active[slot] = (req, h, res, fp)
                used += fp
            if not active and pending and realtime:
                wait = pending[0].arrival - (clock() - t0)
                if wait > 0 and idle is not None:
                    idle(min(wait, 0.002))
            elif idle is not None and active:
                idle(0.0)
        return [results[r.rid] for r in sorted(requests, key=lambda r: r.rid)]


def latency_report(results: List[Result]) -> Dict[str, Any]:
    ok = [r for r in results if r.error is None]
    if not ok:
        return {"completed": 0}
    lat = np.array([r.latency for r in ok])
    span = max(r.done for r in ok) - min(r.arrival for r in ok)
    rep = {"completed": len(ok), "rejected": len(results) - len(ok),
           "req_per_s": len(ok) / span if span > 0 else math.inf,
           "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
           "mean_ms": float(lat.mean() * 1e3), "span_s": span, "per_kind": {}}
    for k in sorted({r.kind for r in ok}):
        lk = np.array([r.latency for r in ok if r.kind == k])
        rep["per_kind"][k] = {"n": int(lk.size), "p50_ms": float(np.percentile(lk, 50) * 1e3),
                              "p99_ms": float(np.percentile(lk, 99) * 1e3)}
    return rep


class GpuExecutor:
    """Runs FDK / FBP plans on one GPU.  Each slot has its own stream, device input
    buffer, plan workspace, pinned host output buffer and (optionally) one captured
    graph per geometry class.

    Classes listed in `chunked` (cone FDK plans; value = number of view ranges) are run
    with ConeFDK.run_chunked instead: the slot's copy stream uploads view range k + 1 while
    its compute stream filters and backprojects range k, so a large request's upload
    overlaps its own reconstruction (eager launches; the chunking is fixed per class, so
    a request's result is the same alone and in any mix)."""

    def __init__(self, plans: Dict[str, Any], slots: int, use_graphs: bool = True, keep_outputs: bool = False,
                 device="cuda", chunked: Dict[str, int] | None = None):
        import torch

        self.torch = torch
        self.plans = plans
        self.device = torch.device(device)
        self.use_graphs = use_graphs
        self.keep = keep_outputs
        self.chunked = dict(chunked or {})
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(slots)]
        self.copy_streams = [torch.cuda.Stream(device=self.device) for _ in range(slots)] if self.chunked else []
        self._slot: Dict[tuple, Dict[str, Any]] = {}
        self._fp = {k: self._estimate(k) for k in plans}

    def _in_shape(self, kind):
        p = self.plans[kind]
        g = p.g
        if hasattr(g, "detector_shape") and isinstance(g.detector_shape, tuple):
            return (1, g.n_views) + tuple(g.detector_shape)
        return (1, g.n_views, g.detector_shape)

    def _out_shape(self, kind):
        return (1,) + tuple(self.plans[kind].g.volume_shape)

    def _estimate(self, kind) -> int:
        p = self.plans[kind]
        nin = int(np.prod(self._in_shape(kind)))
        nout = int(np.prod(self._out_shape(kind)))
        e = p.storage.itemsize
        ws = 8 if getattr(p, "filter_impl", "gemm") == "gemm" else 0   # GEMM filter workspace: 2 x f32
        if kind in self.chunked:
            # two range buffers + filtered range + filter workspace + detector cells of one range + output
            nk = -(-nin // self.chunked[kind])
            return nk * (2 * e + e + ws + 4 * e) + nout * 4
        # input + filtered + filter workspace + detector cells (4 x) + output
        return nin * (e + e + ws + 4 * e) + nout * 4

    def footprint(self, kind) -> int:
        return self._fp[kind]

    def _buffers(self, slot, kind):
        key = (slot, kind)
        b = self._slot.get(key)
        if b is not None:
            return b
        torch = self.torch
        p = self.plans[kind]
        if kind in self.chunked:
            ranges = p.chunk_bounds(self.chunked[kind])
            kmax = max(hi - lo for lo, hi in ranges)
            shape = self._in_shape(kind)
            b = {"bufs": [torch.zeros((1, kmax) + tuple(shape[2:]), dtype=p.storage, device=self.device)
                          for _ in range(2)],
                 "out": torch.empty(self._out_shape(kind), dtype=torch.float32, device=self.device),
                 "host_out": torch.empty(self._out_shape(kind), dtype=torch.float32, pin_memory=True)}
            torch.cuda.synchronize(self.device)
            self._slot[key] = b
            return b
        b = {"in": torch.zeros(self._in_shape(kind), dtype=p.storage, device=self.device),
             "host_out": torch.empty(self._out_shape(kind), dtype=torch.float32, pin_memory=True)}
        if self.use_graphs:
            with torch.cuda.stream(self.streams[slot]):
                b["graph"], b["static_in"], b["static_out"] = p.capture(b["in"])
        torch.cuda.synchronize(self.device)
        self._slot[key] = b
        return b

    def warm(self, kinds=None):
        """Allocate buffers and capture graphs for every (slot, class) up front; chunked
        classes run one filter + view-range backprojection on their (zero) buffers."""
        torch = self.torch
        for s in range(len(self.streams)):
            for k in (kinds or self.plans):
                b = self._buffers(s, k)
                if k in self.chunked:
                    p = self.plans[k]
                    lo, hi = p.chunk_bounds(self.chunked[k])[0]
                    with torch.cuda.stream(self.streams[s]):
                        p.backproject_views(p.filter(b["bufs"][0][:, :hi - lo]), lo, 2, b["out"])
        torch.cuda.synchronize(self.device)

    def start(self, slot, req):
        torch = self.torch
        b = self._buffers(slot, req.kind)
        p = self.plans[req.kind]
        st = self.streams[slot]
        with torch.cuda.stream(st):
            if req.kind in self.chunked:
                # the last view range is backprojected in z slabs, each downloaded while the next is computed
                p.run_chunked(req.payload, b["out"], b["bufs"], self.copy_streams[slot], self.chunked[req.kind],
                              host_out=b["host_out"])
                ev = torch.cuda.Event()
                ev.record(st)
                return (ev, b)
            if self.use_graphs:
                b["static_in"].copy_(req.payload, non_blocking=True)
                b["graph"].replay()
                out = b["static_out"]
            else:
                b["in"].copy_(req.payload, non_blocking=True)
                out = p(b["in"])
            b["host_out"].copy_(out, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        return (ev, b)

    def poll(self, h) -> bool:
        return h[0].query()

    def finish(self, h):
        h[0].synchronize()
        return h[1]["host_out"].clone() if self.keep else None
