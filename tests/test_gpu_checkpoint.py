"""Exact continuation through CanonicalCheckpoint on the MI355X engine
(lib/CorePotts/test/test_phase8_persistence.jl:97-118): an integrator captured at MCS 2,
restored (trackers rebuilt from ownership on the host, state re-uploaded, MCS counter set)
and run 3 more MCS ends bit-exactly where the uninterrupted 5-MCS run ends, with the same
report. The same holds across implementations (a checkpoint of the oracle continued on the
GPU, and a GPU checkpoint continued by the oracle) and for batch replicas restored into a
running batch at their checkpoint MCS."""
import numpy as np
import pytest

from corepotts import ReplicaBatch, workloads
from corepotts import checkpoint as ck
from corepotts.native import ScientificPottsIntegrator
from corepotts.spec import ALG_CHECKERBOARD, ALG_LOTTERY, build_spec, state_arrays
from oracle import bridge
from test_checkpoint import ALGORITHMS, phase8_model, tissue_model

pytestmark = pytest.mark.gpu

FIELDS = ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures", "active",
          "cell_types", "target_volume", "volume_strength", "target_boundary", "boundary_strength")


def _assert_same(observed, expected):
    for key in FIELDS:
        np.testing.assert_array_equal(observed[key], expected[key], err_msg=key)


def _models():
    out = []
    for alg, name in zip(ALGORITHMS, ("checkerboard", "lottery")):
        out.append(("phase8", alg, phase8_model(alg)))
        out.append(("tissue", alg, tissue_model(alg)))
        compiled, proposal, components, algv, seed = workloads.scaled_2d(96, algorithm=name, seed=4321)
        out.append(("scaled96", alg, (compiled, build_spec(compiled, proposal, components, algv, seed))))
    return out


MODELS = _models()


@pytest.mark.parametrize("case", MODELS, ids=[f"{n}-{type(a).__name__}" for n, a, _ in MODELS])
def test_gpu_exact_continuation(case):
    name, _, (compiled, spec) = case
    print(f"checkpoint continuation {name} algorithm={spec.algorithm} dims={spec.dims}", flush=True)
    uninterrupted = ScientificPottsIntegrator(compiled, spec)
    root = ck.capture_checkpoint(uninterrupted)
    assert root.mcs == 0 and root.warnings == ()
    uninterrupted.step(2)
    cp = ck.capture_checkpoint(uninterrupted, ancestry=root)
    assert cp.initial_state_fingerprint == root.initial_state_fingerprint
    uninterrupted.step(3)
    expected = uninterrupted.download()

    resumed = ck.restore_checkpoint(cp, uninterrupted)
    assert resumed.mcs == 2
    resumed.step(3)
    assert resumed.mcs == uninterrupted.mcs == 5
    _assert_same(resumed.download(), expected)
    assert resumed.current_mcs_report() == uninterrupted.current_mcs_report()

    # the oracle continues the GPU checkpoint to the same state
    restored = ck.checkpoint_compiled_state(cp, compiled)
    sim = bridge.OracleSim(spec, state_arrays(restored, spec))
    sim.mcs = cp.mcs
    sim.step(3)
    _assert_same(sim.store(), expected)
    assert sim.report() == tuple(uninterrupted.current_mcs_report())
    resumed.close()
    uninterrupted.close()


@pytest.mark.parametrize("alg", ALGORITHMS, ids=["checkerboard", "lottery"])
def test_gpu_continues_oracle_checkpoint(alg):
    compiled, spec = tissue_model(alg, seed=0xC0FFEE)
    print(f"oracle checkpoint -> GPU algorithm={spec.algorithm}", flush=True)
    sim = bridge.OracleSim(spec, state_arrays(compiled, spec))
    sim.step(4)
    cp = ck.capture_arrays(sim.store(), compiled, spec, sim.mcs, spec.seed)
    sim.step(6)
    prototype = ScientificPottsIntegrator(compiled, spec)
    resumed = ck.restore_checkpoint(cp, prototype)
    resumed.step(6)
    _assert_same(resumed.download(), sim.store())
    assert tuple(resumed.current_mcs_report()) == sim.report()
    resumed.close()
    prototype.close()


@pytest.mark.parametrize("alg", ALGORITHMS, ids=["checkerboard", "lottery"])
def test_batch_replica_restored_mid_run(alg):
    """Replicas checkpointed out of one batch and restored into another that is already
    running other replicas continue exactly."""
    algorithm = ALG_CHECKERBOARD if alg is ALGORITHMS[0] else ALG_LOTTERY
    models = [tissue_model(alg, seed=s) for s in (11, 12)] + [phase8_model(alg)]
    print(f"batch checkpoint restore algorithm={algorithm}", flush=True)
    first = ReplicaBatch(algorithm)
    slots = [first.add(c, s, target_mcs=3) for c, s in models]
    first.step(3)
    cps = [ck.capture_batch_checkpoint(first, slot) for slot in slots]
    assert [cp.mcs for cp in cps] == [3, 3, 3]
    first.close()

    second = ReplicaBatch(algorithm)
    busy = second.add(*tissue_model(alg, seed=99), target_mcs=10)
    second.step(2)
    restored = [ck.restore_into_batch(cp, second, s, c, target_mcs=8) for cp, (c, s) in zip(cps, models)]
    second.step(10)
    done = second.progress()
    assert [done[s] for s in restored] == [8, 8, 8] and done[busy] == 10
    for slot, (compiled, spec) in zip(restored, models):
        sim = bridge.OracleSim(spec, state_arrays(compiled, spec))
        sim.step(8)
        out = second.download(slot)
        _assert_same(out, sim.store())
        assert tuple(second.report(slot)) == sim.report()
    second.close()
