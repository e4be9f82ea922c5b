"""Scheduler logic on the CPU with a fake executor and a fake clock: continuous
admission, HBM-budget refusal, FIFO order, overlap, latency accounting."""
import numpy as np
import pytest

from pyronn_layers.serving import Request, Scheduler, latency_report, poisson_arrivals


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-4)


class FakeExecutor:
    """Each request takes `dur[kind]` seconds of fake-clock time after it starts."""

    def __init__(self, clock, dur, fp):
        self.clock, self.dur, self.fp = clock, dur, fp
        self.running = 0
        self.max_running = 0
        self.starts = []

    def footprint(self, kind):
        return self.fp[kind]

    def start(self, slot, req):
        self.running += 1
        self.max_running = max(self.max_running, self.running)
        self.starts.append((req.rid, slot, self.clock()))
        return {"end": self.clock() + self.dur[req.kind], "rid": req.rid}

    def poll(self, h):
        return self.clock() >= h["end"]

    def finish(self, h):
        self.running -= 1
        return h["rid"] * 10


def test_requests_overlap_and_are_admitted_as_slots_free():
    clk = FakeClock()
    ex = FakeExecutor(clk, {"a": 1.0}, {"a": 10})
    reqs = [Request(i, "a", 0.1 * i) for i in range(6)]
    res = Scheduler(ex, slots=2, hbm_budget=100).run(reqs, clock=clk, idle=clk.sleep)
    assert ex.max_running == 2
    assert all(r.error is None for r in res)
    assert [r.output for r in res] == [i * 10 for i in range(6)]
    # request 2 waits for a slot: admitted when request 0 finishes (~1.0 s)
    assert res[2].admitted == pytest.approx(1.0, abs=2e-3)
    rep = latency_report(res)
    assert rep["completed"] == 6 and rep["p99_ms"] >= rep["p50_ms"] > 0


def test_hbm_budget_limits_concurrency_and_rejects_oversized():
    clk = FakeClock()
    ex = FakeExecutor(clk, {"big": 1.0, "small": 0.1, "huge": 0.1}, {"big": 60, "small": 10, "huge": 1000})
    reqs = [Request(0, "big", 0.0), Request(1, "big", 0.0), Request(2, "small", 0.0), Request(3, "huge", 0.0),
            Request(4, "small", 0.05)]
    res = Scheduler(ex, slots=4, hbm_budget=100).run(reqs, clock=clk, idle=clk.sleep)
    assert res[3].error is not None and "budget" in res[3].error
    # second big request cannot start before the first finishes (60 + 60 > 100)
    assert res[1].admitted >= res[0].done - 1e-9
    # FIFO: the small request queued behind the blocked big one does not overtake it
    assert res[2].admitted >= res[1].admitted
    assert all(r.error is None for r in res if r.rid != 3)


def test_poisson_arrivals_are_seeded():
    a = poisson_arrivals(7, 10.0, 50, {"x": 0.4, "y": 0.6})
    b = poisson_arrivals(7, 10.0, 50, {"x": 0.4, "y": 0.6})
    assert [(r.kind, r.arrival) for r in a] == [(r.kind, r.arrival) for r in b]
    assert np.all(np.diff([r.arrival for r in a]) > 0)
    assert 2.0 < a[-1].arrival < 10.0


def test_executor_refusal_frees_the_slot():
    clk = FakeClock()

    class Refusing(FakeExecutor):
        def start(self, slot, req):
            if req.rid == 0:
                raise RuntimeError("no memory")
            return super().start(slot, req)

    ex = Refusing(clk, {"a": 0.5}, {"a": 1})
    res = Scheduler(ex, slots=1, hbm_budget=10).run([Request(0, "a", 0.0), Request(1, "a", 0.0)], clock=clk,
                                                    idle=clk.sleep)
    assert res[0].error == "no memory" and res[1].error is None and res[1].output == 10


class GroupingExecutor(FakeExecutor):
    """Kinds 'g1'/'g2' share a group key; a group takes the longest member's time."""

    def __init__(self, clock, dur, fp):
        super().__init__(clock, dur, fp)
        self.groups = []

    def group_key(self, kind):
        return "g" if kind.startswith("g") else None

    def start_group(self, slot, reqs):
        self.groups.append([r.rid for r in reqs])
        self.running += 1
        self.max_running = max(self.max_running, self.running)
        return {"end": self.clock() + max(self.dur[r.kind] for r in reqs), "rids": [r.rid for r in reqs]}

    def finish(self, h):
        self.running -= 1
        if "rids" in h:
            return [rid * 10 for rid in h["rids"]]
        return h["rid"] * 10


def test_grouping_takes_a_contiguous_run_of_arrived_requests():
    clk = FakeClock()
    ex = GroupingExecutor(clk, {"g1": 1.0, "g2": 0.5, "x": 0.2}, {"g1": 10, "g2": 10, "x": 10})
    reqs = [Request(0, "g1", 0.0), Request(1, "g2", 0.0), Request(2, "g1", 0.0), Request(3, "x", 0.0),
            Request(4, "g1", 0.0), Request(5, "g2", 5.0)]
    res = Scheduler(ex, slots=2, hbm_budget=1000, group_max=3).run(reqs, clock=clk, idle=clk.sleep)
    assert all(r.error is None for r in res)
    assert [r.output for r in res] == [i * 10 for i in range(6)]
    # 0, 1, 2 start together (the run stops at the 'x' request, no overtaking); 4 runs alone
    # behind 'x'; 5 has not arrived when 4 starts
    assert ex.groups[0] == [0, 1, 2]
    assert [0, 1, 2] == [r.rid for r in res if r.slot == res[0].slot and r.admitted == res[0].admitted]
    assert res[4].admitted >= res[3].admitted
    assert res[5].admitted >= 5.0


def test_grouping_respects_the_budget_and_group_max():
    clk = FakeClock()
    ex = GroupingExecutor(clk, {"g1": 1.0}, {"g1": 40})
    reqs = [Request(i, "g1", 0.0) for i in range(5)]
    Scheduler(ex, slots=1, hbm_budget=100, group_max=4).run(reqs, clock=clk, idle=clk.sleep)
    assert ex.groups[0] == [0, 1]             # a third would exceed the 100-byte budget
    assert all(len(g) <= 2 for g in ex.groups)
    ex2 = GroupingExecutor(clk, {"g1": 1.0}, {"g1": 1})
    Scheduler(ex2, slots=1, hbm_budget=100, group_max=2).run([Request(i, "g1", 0.0) for i in range(5)], clock=clk,
                                                             idle=clk.sleep)
    assert [len(g) for g in ex2.groups] == [2, 2]   # the last request runs alone through start()
