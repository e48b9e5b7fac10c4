"""Mechanics on the MI355X engine against the oracle, bit for bit: the reference's one-MCS
mechanical clock (lib/CorePotts/test/test_scientific_mechanics.jl:107-131), tissues with
volume pressure and surface tension under temperature noise (stationary and constitutive
initialization) for CheckerboardSweepCPM and LotteryCPM, a mechanics replica inside a batch
of plain replicas, and exact continuation through a checkpoint that carries the mechanical
state columns."""
import numpy as np
import pytest

from corepotts import (BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain, CellOwner,
                       CheckerboardSweepCPM, ContactRole, FixedMechanicalNoise,
                       FluctuatingSurfaceTension, FluctuatingVolumePressure, LogicalPottsState,
                       LotteryCPM, MediumOwner, MediumTypeTable, ProposalRole,
                       QuadraticBoundaryHamiltonian, QuadraticVolumeHamiltonian, ReplicaBatch,
                       ScientificComponentSet, StationaryMechanicalInitialization, SurfaceRole,
                       UnorderedContactHamiltonian, build_spec, compile_scientific_state,
                       first_shell_relation, init_scientific, moore_relation, required_properties,
                       workloads)
from corepotts import checkpoint as ck
from corepotts.spec import ALG_CHECKERBOARD, ALG_LOTTERY
from helpers import oracle_for
from test_oracle_mechanics import _clock_fixture

pytestmark = pytest.mark.gpu

FIELDS = ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures", "active", "cell_types",
          "target_volume", "volume_strength", "target_boundary", "boundary_strength")


def _same(g, o, mech=("volume_pressure", "surface_tension")):
    for k in FIELDS:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    for k in mech:
        if k in o:
            np.testing.assert_array_equal(g[k], o[k], err_msg=k)


def mechanical_tissue(algorithm, seed=0xA11CE, dims=(40, 36), init=StationaryMechanicalInitialization,
                      which=("pressure", "tension")):
    """Cells of 6 x 6 in medium with volume, Moore contact and boundary energies plus the
    mechanical components; noise = the algorithm temperature."""
    owners = np.full(dims, MediumOwner(1), dtype=np.int64)
    cid = 0
    for x0 in range(2, dims[0] - 6, 8):
        for y0 in range(2, dims[1] - 6, 8):
            cid += 1
            owners[x0:x0 + 6, y0:y0 + 6] = CellOwner(cid)
    surface = first_shell_relation(SurfaceRole, 2)
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    contact = UnorderedContactHamiltonian(np.array([[0.0, 6.0, 9.0], [6.0, 3.0, 7.0], [9.0, 7.0, 4.0]], np.float32),
                                          MediumTypeTable({1: 1}), moore_relation(ContactRole, 2))
    boundary = QuadraticBoundaryHamiltonian(BoundaryEdgeCount(), surface, number_type=np.float32)
    mechanics = []
    if "pressure" in which:
        mechanics.append(FluctuatingVolumePressure(eta=0.8, initialization=init, instance_id=3))
    if "tension" in which:
        mechanics.append(FluctuatingSurfaceTension(BoundaryEdgeCount(), surface, eta=1.3,
                                                   noise=FixedMechanicalNoise(0.5), initialization=init,
                                                   instance_id=5))
    schema = {}
    for c in (volume, boundary, *mechanics):
        schema.update(required_properties(c))
    state = LogicalPottsState(owners, cid + 2, cell_types={c: 2 + c % 2 for c in range(1, cid + 1)},
                              medium_domains=(1,), property_schema=schema)
    state.property_values("target_volume")[:cid] = 30.0
    state.property_values("volume_strength")[:cid] = 0.7
    state.property_values("target_boundary")[:cid] = 20
    state.property_values("boundary_strength")[:cid] = 0.3
    compiled = compile_scientific_state(state, CartesianDomain(dims), BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    components = ScientificComponentSet(energies=(volume, contact, boundary), mechanics=tuple(mechanics))
    return compiled, first_shell_relation(ProposalRole, 2), components, algorithm, seed


ALGS = [(CheckerboardSweepCPM(np.float32(6.0)), "checkerboard"), (LotteryCPM(np.float32(6.0)), "lottery")]


@pytest.mark.parametrize("alg", ALGS, ids=[a[1] for a in ALGS])
def test_gpu_mechanical_clock(alg):
    algorithm = alg[0].__class__(np.float32(5.0))
    compiled, proposal, components = _clock_fixture(algorithm)
    print(f"mechanical clock {alg[1]}", flush=True)
    gpu = init_scientific(compiled, proposal, components, algorithm, seed=0x91)
    sim, _ = oracle_for(compiled, proposal, components, algorithm, 0x91)
    gpu.step(1)
    sim.step(1)
    g, o = gpu.download(), sim.store()
    _same(g, o)
    assert g["volume_pressure"][0] == pytest.approx(24.0 * (1.0 - np.exp(-1.25)), rel=2e-6)
    assert g["surface_tension"][0] == pytest.approx(-12.0 * (1.0 - np.exp(-0.75)), rel=2e-6)
    gpu.close()


@pytest.mark.parametrize("which", [("pressure",), ("tension",), ("pressure", "tension")],
                         ids=["pressure", "tension", "both"])
@pytest.mark.parametrize("alg", ALGS, ids=[a[1] for a in ALGS])
def test_gpu_mechanical_tissue_parity(alg, which):
    model = mechanical_tissue(alg[0], which=which)
    print(f"mechanical tissue {alg[1]} {which}", flush=True)
    gpu = init_scientific(*model[:4], seed=model[4])
    sim, _ = oracle_for(*model)
    _same(gpu.download(), sim.store())          # initialization
    for _ in range(3):
        gpu.step(2)
        sim.step(2)
        _same(gpu.download(), sim.store())
        assert tuple(gpu.current_mcs_report()) == sim.report()
    gpu.close()


def test_gpu_mechanics_in_a_batch():
    print("mechanics batch", flush=True)
    batch = ReplicaBatch(ALG_CHECKERBOARD)
    models = [mechanical_tissue(CheckerboardSweepCPM(np.float32(6.0)), seed=s) for s in (1, 2)]
    models.append(workloads.scaled_2d(50, seed=7))
    slots = [batch.add_model(*m, target_mcs=4 + k) for k, m in enumerate(models)]
    batch.step(10)
    for slot, m, k in zip(slots, models, range(3)):
        sim, _ = oracle_for(*m)
        sim.step(4 + k)
        _same(batch.download(slot), sim.store())
        assert tuple(batch.report(slot)) == sim.report()
    batch.close()


@pytest.mark.parametrize("alg", ALGS, ids=[a[1] for a in ALGS])
def test_gpu_mechanics_checkpoint_continuation(alg):
    model = mechanical_tissue(alg[0], seed=0xBEEF)
    print(f"mechanics checkpoint {alg[1]}", flush=True)
    gpu = init_scientific(*model[:4], seed=model[4])
    gpu.step(3)
    cp = ck.capture_checkpoint(gpu)
    assert "volume_pressure" in [c.key for c in cp.properties]
    gpu.step(4)
    resumed = ck.restore_checkpoint(cp, gpu)
    resumed.step(4)
    _same(resumed.download(), gpu.download())
    assert resumed.current_mcs_report() == gpu.current_mcs_report()
    resumed.close()
    gpu.close()
