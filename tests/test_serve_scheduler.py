"""Host-side tests of the serving scheduler: admission by budget, first-come-first-served
order, continuous batching (jobs join while others run) and completion accounting. A
stand-in batch with a simulated clock replaces the device engine."""
import numpy as np
import pytest

from corepotts.native import CPM_E_NOMEM, EngineError
from corepotts.serve import JobServer, Request, latency_summary


class FakeSpec:
    def __init__(self, n, nbytes):
        self.dims = (n, 1, 1)
        self.mutable_mask = None
        self.nbytes = nbytes


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class FakeBatch:
    """Replicas advance min(n, target - mcs) per step; a step costs 1 ms per running replica."""

    def __init__(self, clock, engine_budget=None):
        self.clock = clock
        self.mcs, self.target, self.bytes = {}, {}, {}
        self.next = 0
        self.engine_budget = engine_budget
        self.log = []

    def reserve(self, spec):
        if self.engine_budget is not None and sum(self.bytes.values()) + spec.nbytes > self.engine_budget:
            raise EngineError(CPM_E_NOMEM, "HBM budget exceeded")
        s = self.next
        self.next += 1
        self.mcs[s], self.target[s], self.bytes[s] = 0, 0, spec.nbytes
        return s

    def load(self, slot, compiled, target_mcs=None):
        self.target[slot] = target_mcs
        self.log.append(("add", compiled, self.clock()))

    def remove(self, slot):
        del self.mcs[slot], self.target[slot], self.bytes[slot]

    def step(self, n):
        for s in self.mcs:
            self.mcs[s] = min(self.target[s], self.mcs[s] + n)
        self.clock.t += 1e-3 * max(len(self.mcs), 1)

    def progress(self):
        return dict(self.mcs)

    def download(self, slot):
        return {"mcs": self.mcs[slot]}

    def report(self, slot):
        return (self.mcs[slot],)


def _requests(specs, arrivals, mcs):
    return [Request(k, a, k, s, m) for k, (s, a, m) in enumerate(zip(specs, arrivals, mcs))]


def _server(clock, budget, chunk=10, engine_budget=None):
    batch = FakeBatch(clock, engine_budget)
    return JobServer(budget, chunk, batch=batch, bytes_of=lambda s: s.nbytes, clock=clock,
                     sleep=clock.sleep), batch


def test_every_job_completes_once_with_its_mcs():
    clock = FakeClock()
    server, batch = _server(clock, budget=100)
    rng = np.random.default_rng(3)
    specs = [FakeSpec(int(rng.integers(10, 50)), int(rng.integers(5, 40))) for _ in range(40)]
    arrivals = np.cumsum(rng.exponential(0.002, size=40))
    mcs = rng.integers(5, 100, size=40)
    done = server.serve(_requests(specs, arrivals, mcs))
    assert sorted(c.key for c in done) == list(range(40))
    for c in done:
        assert c.state["mcs"] == mcs[c.key]
        assert c.attempts == specs[c.key].dims[0] * mcs[c.key]
        assert c.arrival_s <= c.admitted_s <= c.finished_s
    assert server.peak_used <= 100
    assert not batch.mcs  # everything removed


def test_admission_is_first_come_first_served():
    clock = FakeClock()
    server, batch = _server(clock, budget=50)
    specs = [FakeSpec(10, 30), FakeSpec(10, 30), FakeSpec(10, 5), FakeSpec(10, 5)]
    done = server.serve(_requests(specs, [0.0, 0.0, 0.0, 0.0], [20, 20, 20, 20]))
    order = [entry[1] for entry in batch.log]
    assert order == [0, 1, 2, 3]  # small jobs do not overtake the blocked head
    assert len(done) == 4


def test_jobs_join_while_others_run():
    clock = FakeClock()
    server, batch = _server(clock, budget=1000, chunk=5)
    specs = [FakeSpec(10, 10) for _ in range(3)]
    # job 0 runs long; jobs 1 and 2 arrive while it runs and finish before it
    done = server.serve(_requests(specs, [0.0, 0.0105, 0.0205], [400, 20, 20]))
    by = {c.key: c for c in done}
    assert by[1].admitted_s < by[0].finished_s and by[2].admitted_s < by[0].finished_s
    assert by[1].finished_s < by[0].finished_s
    assert server.peak_jobs >= 2
    assert [c.key for c in done][-1] == 0


def test_engine_budget_refusal_waits_for_completions():
    clock = FakeClock()
    # the engine's own accounting is tighter than the scheduler's estimate
    server, batch = _server(clock, budget=100, engine_budget=60)
    specs = [FakeSpec(10, 40), FakeSpec(10, 40), FakeSpec(10, 40)]
    done = server.serve(_requests(specs, [0.0, 0.0, 0.0], [10, 10, 10]))
    assert len(done) == 3
    assert max(sum(1 for c in done if c.admitted_s <= t < c.finished_s)
               for t in [c.admitted_s for c in done]) == 1


def test_job_larger_than_budget_is_rejected():
    clock = FakeClock()
    server, _ = _server(clock, budget=10)
    with pytest.raises(ValueError):
        server.serve(_requests([FakeSpec(10, 11)], [0.0], [5]))


def test_latency_summary():
    clock = FakeClock()
    server, _ = _server(clock, budget=100)
    specs = [FakeSpec(100, 10) for _ in range(10)]
    done = server.serve(_requests(specs, [0.001 * k for k in range(10)], [50] * 10))
    s = latency_summary(done)
    assert s["jobs"] == 10 and s["p50_s"] <= s["p99_s"] <= s["max_s"]
    assert s["attempts"] == 100 * 50 * 10
