"""The configured workloads at depth against oracle digests (tests/golden/deep_digests.json,
written by tests/golden/make_deep_digests.py).

Each compares the BLAKE2b digest of the ownership and tracker arrays and the 12-field report:
* 4096^2 CheckerboardSweepCPM (the headline workload) at MCS 10, 100 and 1000, HIP graphs of
  10-MCS chunks;
* the same model under LotteryCPM at MCS 10 and 100;
* 256^3 cubes (Float32 J, 26-neighbour contact) at MCS 10 and 200;
* the whole 512-replica batched ensemble at MCS 50 in one ReplicaBatch with 50-MCS graphs;
* eight served jobs at configured sizes (2D sides 1160-2019, 3D sides 111-113), 20 MCS each,
  through the continuous-batching JobServer.
"""
import pytest

from corepotts import JobServer, ReplicaBatch, init_scientific, make_request, workloads
from corepotts.spec import ALG_CHECKERBOARD
from helpers import deep_serving_jobs, load_deep_digests, state_digest

pytestmark = pytest.mark.gpu


def _check(label, state, report, expect):
    assert state_digest(state) == expect["digest"], f"{label}: state differs at MCS {expect['mcs']}"
    assert list(report) == expect["report"], f"{label}: report differs at MCS {expect['mcs']}"


def _deep_single(model, expected, chunk):
    gpu = init_scientific(*model)
    gpu.set_graph_chunk(chunk)
    done = 0
    for e in expected:
        gpu.step(e["mcs"] - done)
        done = e["mcs"]
        _check("gpu", gpu.download(), gpu.current_mcs_report(), e)
    gpu.close()


def test_config2_checkerboard_1000_mcs():
    _deep_single(workloads.config2_large(), load_deep_digests()["config2_checkerboard"], 10)


def test_config2_lottery_100_mcs():
    _deep_single(workloads.config2_large(algorithm="lottery"), load_deep_digests()["config2_lottery"], 10)


def test_config3_cubes_200_mcs():
    _deep_single(workloads.cubes_3d(256), load_deep_digests()["config3_checkerboard"], 10)


def test_config4_ensemble_512_replicas():
    expected = load_deep_digests()["config4_ensemble"]
    batch = ReplicaBatch(ALG_CHECKERBOARD, graph_chunk=50)
    slots = {r: batch.add_model(*workloads.ensemble_replica(r), target_mcs=50) for r in range(512)}
    batch.step(50)
    bad = []
    for r, slot in slots.items():
        e = expected[str(r)][0]
        if state_digest(batch.download(slot)) != e["digest"] or list(batch.report(slot)) != e["report"]:
            bad.append(r)
    batch.close()
    assert not bad, f"{len(bad)} replicas differ, first {bad[:8]}"


def test_config5_served_jobs_at_configured_sizes():
    expected = load_deep_digests()["config5_jobs"]
    jobs = deep_serving_jobs()
    reqs = [make_request(k, 0.01 * k, *workloads.serving_model(j), mcs=j.mcs) for k, j in enumerate(jobs)]
    server = JobServer(200 * 10 ** 9, chunk_mcs=10)
    done = server.serve(reqs)
    assert len(done) == len(jobs)
    for c in done:
        _check(f"job {c.key}", c.state, c.report, expected[str(c.key)][0])
    server.batch.close()
