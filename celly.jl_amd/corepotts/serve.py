"""Served simulation jobs: admission by device-memory budget and continuous batching.

A job is a compiled model plus the number of MCS it asks for. Jobs arrive over time; the
server admits the head of the queue as soon as its replica fits the HBM budget
(cpm_replica_bytes), keeps admitted lattices and trackers resident on the device, and advances
every admitted job together in steps of at most `chunk_mcs` MCS (each replica stops at its
own target on the device). Finished jobs are downloaded and removed between steps, which frees
memory for the next admissions, so jobs that arrive while others run join the running batch.

Each job's trajectory depends only on its own state, seed and MCS counter, so its result
equals a solo run of the same model (ScientificPottsIntegrator or the reference's
init_scientific + step!).

Pre-emption (`preempt_after_s`): when the head of the queue has waited that long without
fitting, running jobs are evicted, largest first, until it fits. An evicted job is captured as
a CanonicalCheckpoint at its last completed MCS (persistence/scientific.jl:250-292), removed
(its device memory returns to the budget) and queued again; on re-admission it is restored from
the checkpoint (state, seed, MCS counter: :464-468) and continues exactly where it stopped, so
its result is still the solo run's. A job is evicted at most once and a resumed job never
evicts others, so every job finishes.
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Any, Callable, Dict, List, NamedTuple, Optional

import numpy as np

from .batch import ReplicaBatch, replica_bytes
from .native import CPM_E_NOMEM, EngineError
from .spec import ALG_CHECKERBOARD, ModelSpec, build_spec


class Request(NamedTuple):
    key: object             # caller's job identifier
    arrival_s: float        # arrival time relative to the start of serve()
    compiled: object        # CompiledScientificState
    spec: ModelSpec
    mcs: int                # MCS requested


class Completion(NamedTuple):
    key: object
    arrival_s: float
    admitted_s: float
    finished_s: float
    mcs: int
    attempts: int           # activated attempts over the job (mutable sites x MCS)
    state: Optional[dict]   # downloaded arrays (cpm_download_state order); None if the job failed
    report: Optional[tuple]  # ScientificMCSReport of the job's last MCS; None if the job failed
    preemptions: int = 0    # times the job was evicted and resumed from a checkpoint
    error: Optional[str] = None  # why the job failed (its own engine or validation error)

    @property
    def latency_s(self) -> float:
        return self.finished_s - self.arrival_s


@dataclasses.dataclass
class _Job:
    r: Request
    queued_s: float                   # when the job (re-)entered the queue
    admitted_s: Optional[float] = None  # first admission
    cp: Any = None                    # checkpoint (or device snapshot) of an evicted job
    cp_bytes: int = 0                 # device bytes the evicted job's snapshot holds
    preemptions: int = 0


def make_request(key, arrival_s: float, state, proposal_relation, components, algorithm,
                 seed: int, mcs: int) -> Request:
    """A job in init_scientific's terms (checkerboard.jl:197-219) plus its MCS count."""
    if mcs < 1:
        raise ValueError("a job asks for at least one MCS")
    spec = build_spec(state, proposal_relation, components, algorithm, seed)
    return Request(key, float(arrival_s), state, spec, int(mcs))


class JobServer:
    """Continuous-batching server over one ReplicaBatch.

    `batch` is normally a ReplicaBatch; anything with the same reserve / load / remove /
    step / progress / download / report (and, for pre-emption, checkpoint / resume) methods
    works (the scheduler is tested on the host with a stand-in). `preempt_after_s`: None
    disables pre-emption. `preempt_to`: "host" keeps an evicted job as a canonical checkpoint
    in host memory (its whole device footprint returns to the budget); "device" keeps it as a
    device snapshot (owner words, cell records, medium volumes: about 4 bytes per site plus 64
    per cell stay charged to the budget) and resumes it with device copies only (batch methods
    snapshot / restore_snapshot, snapshot_bytes). A job whose own reserve, load, resume or
    download raises an engine or validation error completes with `error` set (no state or
    report); its slot and budget share are released and the other jobs are not affected."""

    def __init__(self, hbm_budget_bytes: int, chunk_mcs: int = 50, algorithm: int = ALG_CHECKERBOARD,
                 batch=None, bytes_of: Callable[[ModelSpec], int] = replica_bytes,
                 graph_chunk: Optional[int] = None, clock: Callable[[], float] = time.perf_counter,
                 sleep: Callable[[float], None] = time.sleep, shapes=None,
                 preempt_after_s: Optional[float] = None, preempt_to: str = "host"):
        if chunk_mcs < 1:
            raise ValueError("chunk_mcs must be positive")
        self.budget = int(hbm_budget_bytes)
        self.chunk = int(chunk_mcs)
        # steps between membership changes replay HIP graphs of one server step (chunk_mcs MCS)
        # unless graph_chunk says otherwise (0: plain launches)
        gc = self.chunk if graph_chunk is None else int(graph_chunk)
        self.graph_chunk = gc
        self.batch = batch if batch is not None else ReplicaBatch(algorithm, self.budget, gc, shapes)
        self.bytes_of = bytes_of
        self.clock = clock
        self.sleep = sleep
        self.used = 0
        self.peak_used = 0
        self.peak_jobs = 0
        self.steps = 0
        self.preempt_after = None if preempt_after_s is None else float(preempt_after_s)
        if preempt_to not in ("host", "device"):
            raise ValueError("preempt_to is 'host' or 'device'")
        self.preempt_to = preempt_to
        self.preemptions = 0
        self.failures = 0

    def _preempt(self, need: int, running: Dict[int, "_Job"], queue, now: float, size) -> bool:
        """Evict running jobs, largest first, until `need` bytes fit; False (nothing evicted)
        when the evictable jobs cannot make the room. Synchronizes with the running step."""
        deficit = self.used + need - self.budget
        progress = self.batch.progress()
        victims, freed = [], 0
        for slot in sorted(running, key=lambda s: (-size[id(running[s].r)], s)):
            job = running[slot]
            if job.preemptions or job.cp is not None or progress[slot] >= job.r.mcs:
                continue  # evicted before, resumed, or finishing with this step
            victims.append(slot)
            freed += size[id(job.r)]
            if self.preempt_to == "device":
                freed -= self.batch.snapshot_bytes(job.r.spec)
            if freed >= deficit:
                break
        if freed < deficit:
            return False
        for slot in victims:
            job = running.pop(slot)
            if self.preempt_to == "device":
                job.cp = self.batch.snapshot(slot)
                job.cp_bytes = int(job.cp.bytes)
            else:
                job.cp = self.batch.checkpoint(slot)
                job.cp_bytes = 0
            self.batch.remove(slot)
            self.used -= size[id(job.r)] - job.cp_bytes
            job.preemptions += 1
            job.queued_s = now
            self.preemptions += 1
            queue.append(job)
        return True

    def _fail(self, job: "_Job", slot: Optional[int], err: Exception, t0: float, done: list, size):
        """A job whose own reserve, load, resume or download raised: its slot (if any) and budget
        share are released, a held device snapshot is freed, and it completes with the error.
        The batch and the other jobs carry on."""
        if slot is not None:
            try:
                self.batch.remove(slot)
            except Exception:
                pass  # the batch already dropped the slot
            self.used -= size[id(job.r)]
        if job.cp_bytes:
            job.cp.close()
            self.used -= job.cp_bytes
            job.cp, job.cp_bytes = None, 0
        now = self.clock() - t0
        r = job.r
        done.append(Completion(r.key, r.arrival_s, now if job.admitted_s is None else job.admitted_s, now, 0, 0,
                               None, None, job.preemptions, f"{type(err).__name__}: {err}"))
        self.failures += 1

    def serve(self, requests: List[Request]) -> List[Completion]:
        """Runs the stream to completion; returns completions in finishing order.

        Each round: newly admissible jobs are reserved (schedule compile and allocation on the
        host) while the previous step runs on the GPU; then the step is awaited, finished jobs
        are downloaded and removed, the reserved jobs get their states, and the next step of
        at most `chunk_mcs` MCS is enqueued for every resident job."""
        size = {id(r): self.bytes_of(r.spec) for r in requests}
        for r in requests:
            if size[id(r)] > self.budget:
                raise ValueError(f"job {r.key!r} cannot fit the HBM budget on its own")
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival_s))
        queue: collections.deque = collections.deque()
        reserved: List[tuple] = []
        running: Dict[int, _Job] = {}
        done: List[Completion] = []
        inflight = False
        t0 = self.clock()
        while pending or queue or running or reserved:
            now = self.clock() - t0
            while pending and pending[0].arrival_s <= now:
                r = pending.popleft()
                queue.append(_Job(r, r.arrival_s))
            # admission: the head of the queue joins as soon as it fits (first come, first served)
            while queue:
                job = queue[0]
                need = size[id(job.r)]
                if self.used + need > self.budget:
                    if (self.preempt_after is not None and job.cp is None and running
                            and now - job.queued_s >= self.preempt_after
                            and self._preempt(need, running, queue, now, size)):
                        inflight = False  # the eviction synchronized with the step
                        continue
                    break
                try:
                    slot = self.batch.reserve(job.r.spec)
                except (EngineError, ValueError) as e:
                    if isinstance(e, EngineError) and e.code == CPM_E_NOMEM and (running or reserved):
                        break  # the engine's own accounting says wait for a completion
                    queue.popleft()
                    self._fail(job, None, e, t0, done, size)
                    continue
                queue.popleft()
                self.used += need
                if job.admitted_s is None:
                    job.admitted_s = self.clock() - t0
                reserved.append((slot, job))
            if inflight or running:
                progress = self.batch.progress()
                inflight = False
                for slot in sorted(running):
                    job = running[slot]
                    r = job.r
                    if progress[slot] >= r.mcs:
                        try:
                            state = self.batch.download(slot)
                            report = tuple(self.batch.report(slot))
                        except (EngineError, ValueError) as e:
                            del running[slot]
                            self._fail(job, slot, e, t0, done, size)
                            continue
                        finished = self.clock() - t0
                        self.batch.remove(slot)
                        self.used -= size[id(r)]
                        del running[slot]
                        mutable = int(np.count_nonzero(r.spec.mutable_mask)) if r.spec.mutable_mask is not None \
                            else int(np.prod(r.spec.dims))
                        done.append(Completion(r.key, r.arrival_s, job.admitted_s, finished, r.mcs,
                                               mutable * r.mcs, state, report, job.preemptions))
            for slot, job in reserved:
                try:
                    if job.cp is None:
                        self.batch.load(slot, job.r.compiled, target_mcs=job.r.mcs)
                    elif job.cp_bytes:
                        self.batch.restore_snapshot(slot, job.cp, job.r.compiled, target_mcs=job.r.mcs)
                        job.cp.close()
                        self.used -= job.cp_bytes
                        job.cp, job.cp_bytes = None, 0
                    else:
                        self.batch.resume(slot, job.cp, job.r.compiled, target_mcs=job.r.mcs)
                        job.cp = None
                except (EngineError, ValueError) as e:
                    self._fail(job, slot, e, t0, done, size)
                    continue
                running[slot] = job
            reserved = []
            self.peak_used = max(self.peak_used, self.used)
            self.peak_jobs = max(self.peak_jobs, len(running))
            if running:
                self.batch.step(self.chunk)
                self.steps += 1
                inflight = True
            elif pending and not queue:
                wait = pending[0].arrival_s - (self.clock() - t0)
                if wait > 0:
                    self.sleep(wait)
        return done


def latency_summary(done: List[Completion], wall_s: Optional[float] = None) -> dict:
    """p50 / p99 time-to-completion and aggregate attempts/s of a served stream (failed jobs
    are counted, not timed)."""
    failed = sum(1 for c in done if c.error is not None)
    done = [c for c in done if c.error is None]
    if not done:
        return {"jobs": 0, "failed": failed}
    lat = np.array([c.latency_s for c in done])
    first = min(c.arrival_s for c in done)
    last = max(c.finished_s for c in done)
    span = wall_s if wall_s is not None else last - first
    attempts = sum(c.attempts for c in done)
    return {"jobs": len(done), "failed": failed, "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "max_s": float(lat.max()),
            "attempts": attempts, "span_s": span, "attempts_per_s": attempts / span}
