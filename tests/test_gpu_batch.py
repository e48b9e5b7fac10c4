"""Batched replicas and served jobs versus solo runs: every replica of a batch (different
sizes, dimensions, temperatures, contact matrices, seeds and MCS targets; replicas joining
and leaving mid-run) must end bit-exactly where the oracle's solo run of the same model
ends, with the same last-MCS report."""
import dataclasses

import numpy as np
import pytest

from corepotts import (JobServer, ReplicaBatch, ensemble_seed, init_scientific, make_request,
                       latency_summary, replica_bytes, workloads)
from corepotts.spec import ALG_CHECKERBOARD, ALG_LOTTERY, build_spec
from helpers import oracle_for

pytestmark = pytest.mark.gpu

FIELDS = ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures", "active",
          "cell_types")


def _models(algorithm="checkerboard"):
    """Small replicas of uneven shape: 2D squares of several sides (odd sides need more than
    four colours), one with a boundary constraint, and 3D cubes."""
    rng = np.random.default_rng(77)
    out = []
    for k, (L, ndim, side) in enumerate([(23, 2, 10), (40, 2, 10), (31, 2, 7), (16, 3, 4),
                                         (64, 2, 10), (9, 3, 3)]):
        J = workloads.integer_contact_matrix(rng, 2)
        T = np.float32(rng.uniform(2.0, 30.0))
        out.append(workloads.block_cells(L, ndim, side, side, ntypes=2, J=J, temperature=T,
                                         seed=ensemble_seed(99, k + 1), algorithm=algorithm))
    out.append(workloads.scaled_2d(50, algorithm=algorithm, seed=ensemble_seed(99, 50)))
    return out


def _solo(model, mcs):
    compiled, proposal, components, alg, seed = model
    sim, _ = oracle_for(compiled, proposal, components, alg, seed)
    sim.step(mcs)
    return sim.store(), sim.report()


def _check(got: dict, report, want, label):
    state, rep = want
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], state[f], err_msg=f"{label}: {f}")
    assert tuple(report) == rep, f"{label}: report"


@pytest.mark.parametrize("algorithm", ["checkerboard", "lottery"])
@pytest.mark.parametrize("graph_chunk", [0, 4])
def test_batch_replicas_equal_solo_runs(gpu_lib, algorithm, graph_chunk):
    models = _models(algorithm)
    targets = [7, 12, 3, 12, 9, 1, 12]
    alg_id = ALG_CHECKERBOARD if algorithm == "checkerboard" else ALG_LOTTERY
    batch = ReplicaBatch(alg_id, graph_chunk=graph_chunk)
    slots = []
    for model, target in zip(models, targets):
        compiled, proposal, components, alg, seed = model
        slots.append(batch.add_model(compiled, proposal, components, alg, seed, target_mcs=target))
    batch.step(5)
    batch.step(8)  # replicas stop at their own targets
    prog = batch.progress()
    for s, t in zip(slots, targets):
        assert prog[s] == t
    for k, (s, model, t) in enumerate(zip(slots, models, targets)):
        _check(batch.download(s), batch.report(s), _solo(model, t), f"replica {k}")
    batch.close()


def test_batch_membership_changes_mid_run(gpu_lib):
    models = _models()
    batch = ReplicaBatch(ALG_CHECKERBOARD, graph_chunk=2)
    a = batch.add_model(*models[0], target_mcs=10)
    b = batch.add_model(*models[1], target_mcs=4)
    batch.step(4)
    c = batch.add_model(*models[3], target_mcs=6)   # joins at a's MCS 4
    batch.step(2)
    _check(batch.download(b), batch.report(b), _solo(models[1], 4), "b")
    batch.remove(b)
    d = batch.add_model(*models[4], target_mcs=5)   # reuses b's slot
    batch.step(6)
    for slot, model, t in ((a, models[0], 10), (c, models[3], 6), (d, models[4], 5)):
        _check(batch.download(slot), batch.report(slot), _solo(model, t), f"slot {slot}")
    batch.close()


def test_batch_budget_refuses_admission(gpu_lib):
    # resident bytes + the new replica's estimate must stay within the budget
    model = _models()[4]
    need = replica_bytes(build_spec(*model))
    batch = ReplicaBatch(ALG_CHECKERBOARD, hbm_budget_bytes=2 * need - 1)
    batch.add_model(*model, target_mcs=1)
    from corepotts.native import CPM_E_NOMEM, EngineError
    with pytest.raises(EngineError) as err:
        batch.add_model(*model, target_mcs=1)
    assert err.value.code == CPM_E_NOMEM
    assert len(batch.slots) == 1
    batch.step(1)  # the refused add left the batch usable
    assert batch.progress()[batch.slots[0]] == 1
    batch.close()


def test_batch_equals_single_integrator(gpu_lib):
    # the plain single-lattice path and the batch agree on a mid-size replica
    model = workloads.ensemble_replica(3)
    gpu = init_scientific(*model)
    gpu.step(20)
    batch = ReplicaBatch(ALG_CHECKERBOARD, graph_chunk=10)
    s = batch.add_model(*model, target_mcs=20)
    batch.step(20)
    g, b = gpu.download(), batch.download(s)
    for f in FIELDS:
        np.testing.assert_array_equal(g[f], b[f], err_msg=f)
    assert tuple(gpu.current_mcs_report()) == tuple(batch.report(s))
    batch.close()
    gpu.close()


@pytest.mark.parametrize("graph_chunk", [None, 0])
def test_served_jobs_equal_solo_runs(gpu_lib, graph_chunk):
    # None: the server replays HIP graphs of one server step between membership changes
    jobs = workloads.serving_jobs(8, rate_per_s=400.0, min_2d=20, max_2d=70, min_3d=8, max_3d=16)
    reqs, models = [], {}
    for j in jobs:
        model = workloads.serving_model(j)
        models[j.index] = model
        mcs = 3 + j.mcs % 17
        reqs.append(make_request(j.index, j.arrival_s, *model, mcs=mcs))
    sizes = sorted(replica_bytes(r.spec) for r in reqs)
    # a budget of the three largest forces queueing and staggered admissions
    server = JobServer(sum(sizes[-3:]) + (8 << 20), chunk_mcs=3, graph_chunk=graph_chunk)
    done = server.serve(reqs)
    assert sorted(c.key for c in done) == [j.index for j in jobs]
    assert server.peak_jobs >= 2
    for c in done:
        _check(c.state, c.report, _solo(models[c.key], c.mcs), f"job {c.key}")
    s = latency_summary(done)
    assert s["jobs"] == len(jobs) and s["p99_s"] >= s["p50_s"] > 0


@pytest.mark.parametrize("alg", ["checkerboard", "lottery"])
def test_preempted_jobs_equal_solo_runs(gpu_lib, alg):
    """A running job is evicted through a checkpoint when the head of the queue cannot fit the
    budget beside it, and resumed from its checkpointed state, seed and MCS counter once the
    head finishes: every job, evicted or not, still equals the oracle's solo run."""
    models = {0: workloads.scaled_2d(120, algorithm=alg, seed=ensemble_seed(31, 1)),
              1: workloads.block_cells(96, 2, 10, 12, ntypes=2, J=workloads.J_TWO_TYPES,
                                       temperature=np.float32(12.0), seed=ensemble_seed(31, 2), algorithm=alg),
              2: workloads.scaled_2d(60, algorithm=alg, seed=ensemble_seed(31, 3))}
    mcs = {0: 30, 1: 9, 2: 7}
    reqs = [make_request(k, 0.0, *models[k], mcs=mcs[k]) for k in (0, 1, 2)]
    sizes = [replica_bytes(r.spec) for r in reqs]
    # job 0 alone holds most of the budget: jobs 1 and 2 fit only after it leaves
    server = JobServer(max(sizes) + min(sizes[1:]) // 2, chunk_mcs=3, preempt_after_s=0.0,
                       algorithm=ALG_CHECKERBOARD if alg == "checkerboard" else ALG_LOTTERY)
    done = server.serve(reqs)
    by = {c.key: c for c in done}
    assert sorted(by) == [0, 1, 2]
    assert by[0].preemptions == 1 and server.preemptions >= 1
    assert by[1].finished_s < by[0].finished_s
    for c in done:
        _check(c.state, c.report, _solo(models[c.key], c.mcs), f"job {c.key}")


def test_a_bad_job_fails_alone(gpu_lib):
    """Per-job failure isolation: a job whose ownership names a cell beyond its capacity is
    refused at upload (host-side checks, before any device write) and completes with that
    error; the jobs served beside it still equal their solo runs and the budget drains."""
    models = {k: workloads.scaled_2d(L, seed=ensemble_seed(53, k + 1)) for k, L in enumerate((48, 64, 40))}
    reqs = [make_request(k, 0.001 * k, *models[k], mcs=6 + k) for k in (0, 1, 2)]
    good = reqs[1].compiled
    ids = np.array(good.ids, copy=True)
    ids[np.flatnonzero(np.asarray(good.tags) == 1)[0]] = good.capacity + 5
    reqs[1] = reqs[1]._replace(compiled=dataclasses.replace(good, ids=ids))
    server = JobServer(sum(replica_bytes(r.spec) for r in reqs) + (8 << 20), chunk_mcs=3)
    done = server.serve(reqs)
    by = {c.key: c for c in done}
    assert sorted(by) == [0, 1, 2]
    assert by[1].state is None and "capacity" in by[1].error
    assert server.failures == 1 and server.used == 0
    for k in (0, 2):
        assert by[k].error is None
        _check(by[k].state, by[k].report, _solo(models[k], by[k].mcs), f"job {k}")
    assert latency_summary(done)["failed"] == 1


def test_batch_shapes_per_size_class_keep_results(gpu_lib):
    """Tile shapes tuned per size class (replicas of different shapes run in different launch
    groups) change speed only: the batch still equals the oracle's solo runs."""
    from corepotts import BatchShapes
    from corepotts.autotune import size_class, tile_shapes
    rng = np.random.default_rng(5)
    models = []
    for k, L in enumerate([40, 60, 130, 140, 260]):
        J = workloads.integer_contact_matrix(rng, 2)
        models.append(workloads.block_cells(L, 2, 10, 10, ntypes=2, J=J, temperature=np.float32(8.0),
                                            seed=ensemble_seed(7, k + 1)))
    assert sorted({size_class(build_spec(*m)) for m in models}) == [0, 1, 2]
    shapes = BatchShapes()
    tuned = shapes.tune(models, per_class=2, mcs=2)
    assert len(tuned) == 3
    for k, shape in tuned.items():
        assert shape in dict(tile_shapes(k[1], k[2]))
        assert set(shapes.timings[k]) == set(dict(tile_shapes(k[1], k[2])))
    # every shape was also timed on all classes at once
    assert set(shapes.uniform_timings) == {(ALG_CHECKERBOARD, 2, 1)}
    assert set(shapes.uniform_timings[(ALG_CHECKERBOARD, 2, 1)]) == set(dict(tile_shapes(2, 1)))
    shared = BatchShapes()
    shared.tune(models, per_class=2, mcs=2, greedy=False)
    assert len(set(shared.shapes.values())) == 1 and len(shared.shapes) == 3 and not shared.timings
    # force three different shapes so the batch has one launch group per class
    shapes.shapes = {k: s for k, s in zip(sorted(tuned), [7, 0, 12])}
    batch = ReplicaBatch(ALG_CHECKERBOARD, graph_chunk=3, shapes=shapes)
    slots = [batch.add_model(*m, target_mcs=6) for m in models]
    batch.step(6)
    for s, m in zip(slots, models):
        _check(batch.download(s), batch.report(s), _solo(m, 6), f"slot {s}")
    batch.close()
