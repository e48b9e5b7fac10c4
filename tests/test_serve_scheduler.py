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

    def checkpoint(self, slot):
        self.log.append(("checkpoint", slot, self.mcs[slot], self.clock()))
        return {"mcs": self.mcs[slot]}

    def resume(self, slot, cp, compiled, target_mcs=None):
        self.mcs[slot] = cp["mcs"]
        self.target[slot] = target_mcs
        self.log.append(("resume", compiled, self.clock()))

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


def _server(clock, budget, chunk=10, engine_budget=None, preempt_after_s=None):
    batch = FakeBatch(clock, engine_budget)
    return JobServer(budget, chunk, batch=batch, bytes_of=lambda s: s.nbytes, clock=clock,
                     sleep=clock.sleep, preempt_after_s=preempt_after_s), batch


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


def test_preemption_admits_a_blocked_head_and_resumes_exactly():
    """A long job holding most of the budget is evicted through a checkpoint when the head of
    the queue has waited preempt_after_s; it resumes from its checkpointed MCS afterwards and
    still performs exactly the MCS it asked for."""
    clock = FakeClock()
    server, batch = _server(clock, budget=100, chunk=10, preempt_after_s=0.005)
    specs = [FakeSpec(10, 80), FakeSpec(10, 50)]
    done = server.serve(_requests(specs, [0.0, 0.0015], [1000, 30]))
    by = {c.key: c for c in done}
    assert by[0].state["mcs"] == 1000 and by[1].state["mcs"] == 30
    assert by[0].preemptions == 1 and by[1].preemptions == 0 and server.preemptions == 1
    assert by[1].finished_s < by[0].finished_s
    assert by[1].admitted_s - by[1].arrival_s < 0.02  # not held until job 0 finishes
    ck = [e for e in batch.log if e[0] == "checkpoint"]
    rs = [e for e in batch.log if e[0] == "resume"]
    assert len(ck) == 1 and len(rs) == 1 and 0 < ck[0][2] < 1000
    assert server.peak_used <= 100


def test_without_preemption_the_head_waits():
    clock = FakeClock()
    server, _ = _server(clock, budget=100, chunk=10)
    done = server.serve(_requests([FakeSpec(10, 80), FakeSpec(10, 50)], [0.0, 0.0015], [1000, 30]))
    by = {c.key: c for c in done}
    assert by[1].admitted_s >= by[0].finished_s and server.preemptions == 0


def test_preemption_cannot_ping_pong():
    """Evicted jobs are never evicted again and resumed jobs never evict: a stream of large jobs
    that cannot share the budget still finishes, each job evicted at most once."""
    clock = FakeClock()
    server, batch = _server(clock, budget=100, chunk=10, preempt_after_s=0.0)
    specs = [FakeSpec(10, 60) for _ in range(6)]
    mcs = [200, 150, 120, 90, 60, 30]
    done = server.serve(_requests(specs, [0.001 * k for k in range(6)], mcs))
    assert sorted(c.key for c in done) == list(range(6))
    for c in done:
        assert c.state["mcs"] == mcs[c.key]
        assert c.preemptions <= 1
    assert server.peak_used <= 100
    assert not batch.mcs


class FakeSnapshot:
    def __init__(self, mcs, nbytes):
        self.mcs, self.bytes, self.closed = mcs, nbytes, False

    def close(self):
        self.closed = True


class FakeSnapshotBatch(FakeBatch):
    """FakeBatch with device snapshots of a fixed 20 % of a replica's bytes."""

    @staticmethod
    def snapshot_bytes(spec):
        return spec.nbytes // 5

    def snapshot(self, slot):
        self.log.append(("snapshot", slot, self.mcs[slot], self.clock()))
        return FakeSnapshot(self.mcs[slot], self.bytes[slot] // 5)

    def restore_snapshot(self, slot, snap, compiled, target_mcs=None):
        self.mcs[slot] = snap.mcs
        self.target[slot] = target_mcs
        self.log.append(("restore", compiled, self.clock()))


def test_preemption_into_device_snapshots_keeps_their_bytes_charged():
    """preempt_to="device": the evicted job's snapshot stays charged to the budget while it
    waits (so the head is admitted only if the rest fits beside it), the job resumes from the
    snapshot's MCS and performs exactly the MCS it asked for, and the snapshot is released."""
    clock = FakeClock()
    batch = FakeSnapshotBatch(clock)
    server = JobServer(100, chunk_mcs=10, batch=batch, bytes_of=lambda s: s.nbytes, clock=clock,
                       sleep=clock.sleep, preempt_after_s=0.005, preempt_to="device")
    specs = [FakeSpec(10, 80), FakeSpec(10, 50)]
    done = server.serve(_requests(specs, [0.0, 0.0015], [1000, 30]))
    by = {c.key: c for c in done}
    assert by[0].state["mcs"] == 1000 and by[1].state["mcs"] == 30
    assert by[0].preemptions == 1 and server.preemptions == 1
    assert by[1].finished_s < by[0].finished_s
    snaps = [e for e in batch.log if e[0] == "snapshot"]
    restores = [e for e in batch.log if e[0] == "restore"]
    assert len(snaps) == 1 and len(restores) == 1 and 0 < snaps[0][2] < 1000
    assert not [e for e in batch.log if e[0] in ("checkpoint", "resume")]
    assert server.peak_used <= 100 and server.used == 0
    # 80 + 50 > 100, but 80 // 5 + 50 <= 100: the head fits beside the snapshot


def test_device_preemption_that_cannot_make_room_evicts_nothing():
    """A snapshot that would keep too much of the budget frees too little: nothing is evicted
    and the head waits for the running job to finish."""
    clock = FakeClock()
    batch = FakeSnapshotBatch(clock)
    server = JobServer(100, chunk_mcs=10, batch=batch, bytes_of=lambda s: s.nbytes, clock=clock,
                       sleep=clock.sleep, preempt_after_s=0.0, preempt_to="device")
    specs = [FakeSpec(10, 60), FakeSpec(10, 95)]  # 60 // 5 + 95 > 100
    done = server.serve(_requests(specs, [0.0, 0.0015], [200, 30]))
    by = {c.key: c for c in done}
    assert server.preemptions == 0 and by[1].admitted_s >= by[0].finished_s
    assert server.used == 0


def test_preempt_to_must_name_host_or_device():
    with pytest.raises(ValueError):
        JobServer(100, batch=FakeBatch(FakeClock()), preempt_to="disk")


class FailingBatch(FakeBatch):
    """FakeBatch whose reserve, load or download fails for chosen jobs (by request state)."""

    def __init__(self, clock, bad_reserve=(), bad_load=(), bad_download=()):
        super().__init__(clock)
        self.bad_reserve, self.bad_load, self.bad_download = set(bad_reserve), set(bad_load), set(bad_download)
        self.owner = {}

    def reserve(self, spec):
        if spec.nbytes in self.bad_reserve:
            raise EngineError(-2, "unsupported model")
        return super().reserve(spec)

    def load(self, slot, compiled, target_mcs=None):
        if compiled in self.bad_load:
            self.remove(slot)  # as ReplicaBatch.load does on a failed upload
            raise ValueError("owner identity exceeds the capacity")
        self.owner[slot] = compiled
        super().load(slot, compiled, target_mcs)

    def download(self, slot):
        if self.owner.get(slot) in self.bad_download:
            raise EngineError(-1, "download failed")
        return super().download(slot)


def test_a_failing_job_does_not_poison_the_stream():
    """Per-job failure isolation: jobs whose reserve, load or download raise complete with their
    error; every other job completes with its MCS, and the budget is fully released."""
    clock = FakeClock()
    # request k has compiled state k; reserve fails for the spec of 33 bytes
    specs = [FakeSpec(10, 20), FakeSpec(10, 33), FakeSpec(10, 20), FakeSpec(10, 20), FakeSpec(10, 20)]
    batch = FailingBatch(clock, bad_reserve={33}, bad_load={2}, bad_download={4})
    server = JobServer(100, chunk_mcs=10, batch=batch, bytes_of=lambda s: s.nbytes, clock=clock, sleep=clock.sleep)
    done = server.serve(_requests(specs, [0.0, 0.001, 0.002, 0.003, 0.004], [30, 30, 30, 30, 30]))
    by = {c.key: c for c in done}
    assert sorted(by) == [0, 1, 2, 3, 4]
    for k in (1, 2, 4):
        assert by[k].error is not None and by[k].state is None and by[k].report is None
    assert "unsupported" in by[1].error and "capacity" in by[2].error and "download" in by[4].error
    for k in (0, 3):
        assert by[k].error is None and by[k].state["mcs"] == 30
    assert server.failures == 3 and server.used == 0 and not batch.mcs
    s = latency_summary(done)
    assert s["jobs"] == 2 and s["failed"] == 3
