"""Continuous-batching server for reconstruction (FDK / FBP) and projection requests of
mixed sizes.

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
    payload: Any = None       # executor-specific input (host sinogram or volume)


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
        poll(handle) -> bool, finish(handle) -> output,
    and, for grouping (group_max > 1), group_key(kind) -> key or None and
    start_group(slot, requests) -> handle whose finish() returns one output per request.

    With group_max > 1, a request whose kind has a group key is started together with
    the requests right behind it in the queue that have arrived, share its key and fit
    the budget (at most group_max in all, never overtaking a request of another key):
    one slot runs them as one batched launch."""

    def __init__(self, executor, slots: int, hbm_budget: int, group_max: int = 1):
        if slots < 1:
            raise ValueError("need at least one slot")
        if group_max < 1:
            raise ValueError("group_max must be at least 1")
        self.ex = executor
        self.slots = slots
        self.budget = int(hbm_budget)
        self.group_max = int(group_max)

    def run(self, requests: List[Request], clock: Callable[[], float] = time.perf_counter,
            idle: Callable[[float], None] = None, realtime: bool = True) -> List[Result]:
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        free = list(range(self.slots))
        heapq.heapify(free)
        active = {}            # slot -> (results, handle, footprint, grouped)
        used = 0
        results: Dict[int, Result] = {}
        t0 = clock()
        grouping = self.group_max > 1 and hasattr(self.ex, "group_key")
        while pending or active:
            now = clock() - t0
            # retire finished work, in slot order
            for slot in sorted(active):
                ress, h, fp, grouped = active[slot]
                if self.ex.poll(h):
                    out = self.ex.finish(h)
                    done = clock() - t0
                    for k, res in enumerate(ress):
                        res.output = out[k] if grouped else out
                        res.done = done
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
                group = [req]
                key = self.ex.group_key(req.kind) if grouping else None
                if key is not None:
                    while (len(group) < self.group_max and pending and (pending[0].arrival <= now or not realtime)
                           and self.ex.group_key(pending[0].kind) == key):
                        f2 = self.ex.footprint(pending[0].kind)
                        if used + fp + f2 > self.budget:
                            break
                        group.append(pending.popleft())
                        fp += f2
                slot = heapq.heappop(free)
                admitted = clock() - t0
                ress = [Result(q.rid, q.kind, q.arrival, admitted=admitted, slot=slot) for q in group]
                try:
                    h = self.ex.start(slot, req) if len(group) == 1 else self.ex.start_group(slot, group)
                except Exception as e:  # executor refused: report, free the slot
                    for res in ress:
                        res.error = str(e)
                        results[res.rid] = res
                    heapq.heappush(free, slot)
                    continue
                for res in ress:
                    results[res.rid] = res
                active[slot] = (ress, h, fp, len(group) > 1)
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
    """Runs plans on one GPU: FDK / FBP reconstructions (fdk.py) and forward projections
    (projectors.py), one request kind per plan, its geometry resident on the device.
    Each slot has its own stream, device input buffer, plan workspace, pinned host
    output buffer and (optionally) one captured graph per request kind.

    Classes listed in `chunked` (cone FDK plans; value = number of view ranges) are run
    with ConeFDK.run_chunked instead: the slot's copy stream uploads view range k + 1 while
    its compute stream filters and backprojects range k, so a large request's upload
    overlaps its own reconstruction (eager launches; the chunking is fixed per class, so
    a request's result is the same alone and in any mix)."""

    def __init__(self, plans: Dict[str, Any], slots: int, use_graphs: bool = True, keep_outputs: bool = False,
                 device="cuda", chunked: Dict[str, int] | None = None, group_kinds=()):
        import torch

        self.torch = torch
        self.plans = plans
        self.device = torch.device(device)
        self.use_graphs = use_graphs
        self.keep = keep_outputs
        self.chunked = dict(chunked or {})
        # cone FDK kinds whose queued requests may share one batched backprojection launch
        self.group_kinds = set(group_kinds) - set(self.chunked)
        self._host_pool: Dict[tuple, list] = {}
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(slots)]
        self.copy_streams = [torch.cuda.Stream(device=self.device) for _ in range(slots)] if self.chunked else []
        self._slot: Dict[tuple, Dict[str, Any]] = {}
        self._fp = {k: self._estimate(k) for k in plans}

    def _in_shape(self, kind):
        return tuple(self.plans[kind].input_shape)

    def _out_shape(self, kind):
        return tuple(self.plans[kind].output_shape)

    def _estimate(self, kind) -> int:
        p = self.plans[kind]
        if kind in self.chunked:
            # two range buffers + filtered range + filter workspace + detector cells of one range + output
            nin = int(np.prod(self._in_shape(kind)))
            nout = int(np.prod(self._out_shape(kind)))
            e = p.storage.itemsize
            ws = 8 if getattr(p, "filter_impl", "gemm") == "gemm" else 0   # GEMM filter workspace: 2 x f32
            nk = -(-nin // self.chunked[kind])
            return nk * (2 * e + e + ws + 4 * e) + nout * 4
        # the plan's own count: request buffers, workspaces and output
        return int(p.footprint())

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

    def group_key(self, kind):
        """Requests of group kinds with the same storage dtype can share a launch."""
        if kind not in self.group_kinds:
            return None
        return ("cone_fdk", self.plans[kind].storage)

    def start_group(self, slot, reqs):
        """Upload and filter each request, then backproject all of them in ONE batched
        launch (pyronn_cone_backprojection3d_batched over their geometries, each plan's
        device-resident matrices and FDK scale).  The batched launch gives every volume
        bit for bit what the plan's own backprojection gives, so a request's result does
        not depend on its group."""
        from . import batched

        torch = self.torch
        st = self.streams[slot]
        seen: Dict[str, int] = {}
        hosts = []
        with torch.cuda.stream(st):
            qs, plans = [], []
            for r in reqs:
                p = self.plans[r.kind]
                x = torch.empty(p.input_shape, dtype=p.storage, device=self.device)
                x.copy_(r.payload, non_blocking=True)
                qs.append(p.filter(x)[0])
                plans.append(p)
            vols = batched.cone_backprojection3d_batched(qs, [p.g for p in plans], device_geometry=plans)
            for r, v in zip(reqs, vols):
                j = seen.get(r.kind, 0)
                seen[r.kind] = j + 1
                pool = self._host_pool.setdefault((slot, r.kind), [])
                while len(pool) <= j:
                    pool.append(torch.empty(self._out_shape(r.kind), dtype=torch.float32, pin_memory=True))
                pool[j][0].copy_(v, non_blocking=True)
                hosts.append(pool[j])
            ev = torch.cuda.Event()
            ev.record(st)
        return (ev, {"group": hosts})

    def poll(self, h) -> bool:
        return h[0].query()

    def finish(self, h):
        h[0].synchronize()
        if "group" in h[1]:
            return [x.clone() if self.keep else None for x in h[1]["group"]]
        return h[1]["host_out"].clone() if self.keep else None
