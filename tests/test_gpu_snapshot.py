"""Device-resident snapshots (cpm_snapshot_*): a replica's state at a completed MCS kept in
device memory. Restoring one continues exactly (state, seed, MCS counter), in the same or
another context and in a batch slot; its download equals the replica's download at capture,
and its canonical checkpoint equals the host capture at that MCS. Served jobs pre-empted into
device snapshots still equal their solo oracle runs."""
import numpy as np
import pytest

from corepotts import JobServer, ReplicaBatch, ensemble_seed, init_scientific, make_request, workloads
from corepotts import checkpoint as ck
from corepotts.batch import replica_bytes
from corepotts.native import EngineError
from corepotts.spec import ALG_CHECKERBOARD, ALG_LOTTERY, build_spec
from helpers import oracle_for
from test_gpu_mechanics import ALGS, mechanical_tissue

pytestmark = pytest.mark.gpu

FIELDS = ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures", "active", "cell_types",
          "target_volume", "volume_strength", "target_boundary", "boundary_strength")
MECH = ("volume_pressure", "surface_tension")


def _same(a, b, label=""):
    for k in FIELDS + MECH:
        if k in a or k in b:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{label}{k}")


@pytest.mark.parametrize("alg", ["checkerboard", "lottery"])
def test_snapshot_restore_continues_exactly(gpu_lib, alg):
    model = workloads.scaled_2d(96, algorithm=alg, seed=0x5A17)
    gpu = init_scientific(*model)
    gpu.step(7)
    at7 = gpu.download()
    snap = gpu.snapshot()
    assert snap.mcs == 7 and snap.seed == model[4]
    assert snap.bytes == ReplicaBatch.snapshot_bytes(build_spec(*model))
    _same(snap.download(), at7, "snapshot download ")
    gpu.step(5)
    after = gpu.download(), tuple(gpu.current_mcs_report())
    gpu.restore_snapshot(snap)
    assert gpu.mcs == 7
    _same(gpu.download(), at7, "restored ")
    gpu.step(5)
    _same(gpu.download(), after[0], "continued ")
    assert tuple(gpu.current_mcs_report()) == after[1]
    # into another context of the same model, fresh at MCS 0
    other = init_scientific(*model)
    other.restore_snapshot(snap)
    other.step(5)
    _same(other.download(), after[0], "other context ")
    assert tuple(other.current_mcs_report()) == after[1]
    snap.close()
    other.close()
    gpu.close()


def test_snapshot_checkpoint_equals_host_capture(gpu_lib):
    compiled, proposal, components, algorithm, seed = workloads.scaled_2d(64, seed=0xC0DE)
    gpu = init_scientific(compiled, proposal, components, algorithm, seed=seed)
    gpu.step(6)
    host = ck.capture_checkpoint(gpu)
    snap = gpu.snapshot()
    dev = snap.to_checkpoint(compiled)
    assert dev.state_fingerprint == host.state_fingerprint
    assert dev.checksum == host.checksum
    ck.validate_checkpoint(dev)
    snap.close()
    gpu.close()


def test_snapshot_of_another_model_is_rejected(gpu_lib):
    a = init_scientific(*workloads.scaled_2d(64, seed=1))
    b = init_scientific(*workloads.scaled_2d(80, seed=1))
    a.step(1)
    snap = a.snapshot()
    with pytest.raises(EngineError):
        b.restore_snapshot(snap)
    snap.close()
    a.close()
    b.close()


@pytest.mark.parametrize("alg", ALGS, ids=[a[1] for a in ALGS])
def test_snapshot_keeps_mechanical_state(gpu_lib, alg):
    model = mechanical_tissue(alg[0], seed=0xFACE)
    gpu = init_scientific(*model[:4], seed=model[4])
    gpu.step(3)
    snap = gpu.snapshot()
    gpu.step(4)
    after = gpu.download(), tuple(gpu.current_mcs_report())
    other = init_scientific(*model[:4], seed=model[4])
    other.restore_snapshot(snap)
    other.step(4)
    _same(other.download(), after[0], "mechanics ")
    assert tuple(other.current_mcs_report()) == after[1]
    snap.close()
    other.close()
    gpu.close()


def test_batch_snapshot_into_a_new_slot(gpu_lib):
    models = [workloads.block_cells(L, 2, 10, 12, ntypes=2, J=workloads.J_TWO_TYPES, temperature=np.float32(9.0),
                                    seed=ensemble_seed(77, k + 1)) for k, L in enumerate((48, 61, 90))]
    ref = ReplicaBatch(ALG_CHECKERBOARD, graph_chunk=0)
    slots = [ref.add_model(*m, target_mcs=9) for m in models]
    ref.step(9)
    expect = {s: (ref.download(s), tuple(ref.report(s))) for s in slots}
    ref.close()

    batch = ReplicaBatch(ALG_CHECKERBOARD, graph_chunk=0)
    slots = [batch.add_model(*m, target_mcs=4) for m in models]
    batch.step(4)
    snap = batch.snapshot(slots[1])
    assert snap.mcs == 4
    batch.remove(slots[1])
    for s in (slots[0], slots[2]):
        batch.set_target(s, 9)
    batch.step(2)  # the others move on while the replica is out
    spec = build_spec(*models[1])
    slot = batch.reserve(spec)
    batch.restore_snapshot(slot, snap, models[1][0], target_mcs=9)
    snap.close()
    batch.step(7)
    for old, new in ((slots[0], slots[0]), (slots[1], slot), (slots[2], slots[2])):
        _same(batch.download(new), expect[old][0], f"slot {new} ")
        assert tuple(batch.report(new)) == expect[old][1]
    batch.close()


def _solo(model, mcs):
    sim, _ = oracle_for(*model)
    sim.step(mcs)
    return sim.store(), sim.report()


@pytest.mark.parametrize("alg", ["checkerboard", "lottery"])
def test_jobs_preempted_into_device_snapshots_equal_solo_runs(gpu_lib, alg):
    """As the host pre-emption test, with evicted jobs kept as device snapshots: the snapshot's
    bytes stay charged to the budget while the job waits, and every job equals its solo run."""
    models = {0: workloads.scaled_2d(120, algorithm=alg, seed=ensemble_seed(41, 1)),
              1: workloads.block_cells(96, 2, 10, 12, ntypes=2, J=workloads.J_TWO_TYPES,
                                       temperature=np.float32(12.0), seed=ensemble_seed(41, 2), algorithm=alg),
              2: workloads.scaled_2d(60, algorithm=alg, seed=ensemble_seed(41, 3))}
    mcs = {0: 30, 1: 9, 2: 7}
    reqs = [make_request(k, 0.0, *models[k], mcs=mcs[k]) for k in (0, 1, 2)]
    sizes = [replica_bytes(r.spec) for r in reqs]
    budget = max(sizes) + min(sizes[1:]) // 2 + ReplicaBatch.snapshot_bytes(reqs[0].spec)
    server = JobServer(budget, chunk_mcs=3, preempt_after_s=0.0, preempt_to="device",
                       algorithm=ALG_CHECKERBOARD if alg == "checkerboard" else ALG_LOTTERY)
    done = server.serve(reqs)
    by = {c.key: c for c in done}
    assert sorted(by) == [0, 1, 2]
    assert by[0].preemptions == 1 and server.preemptions >= 1
    assert server.used == 0
    for c in done:
        o, rep = _solo(models[c.key], c.mcs)
        for k in ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures"):
            np.testing.assert_array_equal(c.state[k], o[k], err_msg=f"job {c.key} {k}")
        assert tuple(c.report) == rep, f"job {c.key} report"
