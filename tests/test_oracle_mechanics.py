"""Mechanics in the oracle, pinned by lib/CorePotts/test/test_scientific_mechanics.jl and
test_rng_semantics.jl: the OU transition's closed form, the mechanical work of one copy,
the one-MCS mechanical clock of every algorithm, the evolution kernel's moments and the
Box-Muller normal's moments; plus the normal against its formula from the Philox words.
Float32 exp/log/cos are evaluated in double precision and rounded once (parity with
Julia's Float32 functions unpinned, within 1 ulp)."""
import numpy as np
import pytest

from corepotts import (BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain, CellOwner,
                       CheckerboardSweepCPM, FixedMechanicalNoise, FluctuatingSurfaceTension,
                       FluctuatingVolumePressure, LogicalPottsState, LotteryCPM, MediumOwner,
                       PreserveMechanicalInitialization, ProposalRole, ScientificComponentSet,
                       SequentialCPM, StationaryMechanicalInitialization, SurfaceRole, build_spec,
                       compile_scientific_state, first_shell_relation, required_properties,
                       state_arrays)
from oracle import bridge

AUX_EVOLUTION = 9


def _schema(*components):
    out = {}
    for c in components:
        out.update(required_properties(c))
    return out


def test_ou_transition_closed_form():
    # test_scientific_mechanics.jl:55-60
    q = bridge.ou_transition(1.5, 3.0, 2.0, 0.75, 4.0, 0.2, -0.5)
    alpha = np.exp(-0.75 * 0.2)
    expected = alpha * 1.5 + (1 - alpha) * 12.0 - 0.5 * np.sqrt(16.0 * (1 - alpha ** 2))
    assert q == pytest.approx(expected, rel=2e-6)
    assert bridge.ou_transition(1.5, 3.0, 2.0, 0.75, 0.0, 0.0, 7.0) == np.float32(1.5)
    # noise-free relaxation toward the constitutive mean 2 * strength * displacement
    assert bridge.ou_transition(0.0, 3.0, 2.0, 1.0, 0.0, 50.0, 0.0) == pytest.approx(12.0, rel=1e-6)


def test_normal_box_muller_formula_and_moments():
    seed = 0x12345678
    values = []
    for mcs in range(1, 20001):
        z = bridge.normal_box_muller(seed, AUX_EVOLUTION, mcs, 0, 1, 3, 0)
        values.append(z)
        if mcs <= 200:
            w = bridge.rng_words(seed, AUX_EVOLUTION, mcs=mcs, operation=1, kind=2, entity=3)
            u1 = (np.float32(w[0] >> 8) + np.float32(0.5)) * np.float32(2.0 ** -24)
            u2 = (np.float32(w[1] >> 8) + np.float32(0.5)) * np.float32(2.0 ** -24)
            ref = np.sqrt(-2.0 * np.log(np.float64(u1))) * np.cos(2.0 * np.float64(np.float32(np.pi)) * np.float64(u2))
            assert z == pytest.approx(ref, rel=5e-6, abs=5e-6)
    v = np.asarray(values, np.float64)
    # test_rng_semantics.jl:93-107 (moments of the normal draws)
    assert abs(v.mean()) < 0.025
    assert abs(v.var() - 1.0) < 0.04


def _mechanics_pair(noise=0.0, init=PreserveMechanicalInitialization):
    surface = first_shell_relation(SurfaceRole, 2)
    volume = FluctuatingVolumePressure(eta=1.25, noise=FixedMechanicalNoise(noise), initialization=init,
                                       instance_id=1)
    tension = FluctuatingSurfaceTension(BoundaryEdgeCount(), surface, eta=0.75,
                                        noise=FixedMechanicalNoise(noise), initialization=init, instance_id=2)
    return volume, tension, surface


def test_mechanical_work_of_one_copy():
    # test_scientific_mechanics.jl:64-91: cell 1 at the centre of a 3 x 3 medium
    volume, tension, surface = _mechanics_pair()
    owners = np.full((3, 3), MediumOwner(1), dtype=np.int64)
    owners[1, 1] = CellOwner(1)
    state = LogicalPottsState(owners, 1, cell_types={1: 1}, medium_domains=(1,),
                              property_schema=_schema(volume, tension))
    state.property_values("volume_pressure")[0] = 3.0
    state.property_values("surface_tension")[0] = 2.0
    compiled = compile_scientific_state(state, CartesianDomain((3, 3)), BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    spec = build_spec(compiled, first_shell_relation(ProposalRole, 2), ScientificComponentSet(mechanics=(volume, tension)),
                      SequentialCPM(np.float32(1.0)), 0)
    sim = bridge.OracleSim(spec, state_arrays(compiled, spec))
    # CopyProposal(5, 4): recipient site 5 (the cell), donor site 4 = direction (-1, 0) = 1
    outcome, dh, work = sim.copy_evaluation(5, 1)
    assert outcome == 1
    assert dh == 0.0
    # the cell loses its only site: its four outward incidences vanish (losing boundary -4)
    assert work == np.float32(-3.0) + np.float32(2.0) * np.float32(-4.0)


def _clock_fixture(algorithm):
    volume, tension, surface = _mechanics_pair()
    owners = np.full((4, 4), CellOwner(1), dtype=np.int64)
    state = LogicalPottsState(owners, 1, cell_types={1: 1}, medium_domains=(1,),
                              property_schema=_schema(volume, tension))
    state.property_values("target_volume")[0] = 10.0
    state.property_values("volume_strength")[0] = 2.0
    state.property_values("target_boundary")[0] = 4
    state.property_values("boundary_strength")[0] = 1.5
    compiled = compile_scientific_state(state, CartesianDomain((4, 4)), BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    return compiled, first_shell_relation(ProposalRole, 2), ScientificComponentSet(mechanics=(volume, tension))


@pytest.mark.parametrize("algorithm", [SequentialCPM(np.float32(5.0)), CheckerboardSweepCPM(np.float32(5.0)),
                                       LotteryCPM(np.float32(5.0))], ids=["sequential", "checkerboard", "lottery"])
def test_one_mechanical_mcs_across_algorithms(algorithm):
    # test_scientific_mechanics.jl:107-131: the half intervals of one MCS sum to one
    compiled, proposal, components = _clock_fixture(algorithm)
    spec = build_spec(compiled, proposal, components, algorithm, 0x91)
    sim = bridge.OracleSim(spec, state_arrays(compiled, spec))
    sim.initialize_mechanics()
    sim.step(1)
    out = sim.store()
    assert out["volume_pressure"][0] == pytest.approx(24.0 * (1.0 - np.exp(-1.25)), rel=2e-6)
    assert out["surface_tension"][0] == pytest.approx(-12.0 * (1.0 - np.exp(-0.75)), rel=2e-6)


def test_evolution_kernel_moments():
    # test_scientific_mechanics.jl:160-188: 4096 one-site cells at their target volume
    dims = (64, 64)
    n = dims[0] * dims[1]
    component = FluctuatingVolumePressure(eta=1.0, initialization=PreserveMechanicalInitialization, instance_id=11)
    owners = np.arange(1, n + 1, dtype=np.int64).reshape(dims, order="F")
    state = LogicalPottsState(owners, n, cell_types={c: 1 for c in range(1, n + 1)}, medium_domains=(1,),
                              property_schema=required_properties(component))
    state.property_values("target_volume")[:] = 1.0
    state.property_values("volume_strength")[:] = 2.0
    surface = first_shell_relation(SurfaceRole, 2)
    compiled = compile_scientific_state(state, CartesianDomain(dims), BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    spec = build_spec(compiled, first_shell_relation(ProposalRole, 2), ScientificComponentSet(mechanics=(component,)),
                      SequentialCPM(np.float32(3.0)), 0x65766F6C7574696F)
    sim = bridge.OracleSim(spec, state_arrays(compiled, spec))
    sim.advance_mechanics(1, 0, 0, 1.0)
    samples = sim.store()["volume_pressure"].astype(np.float64)
    assert abs(samples.mean()) < 0.2
    assert abs(samples.var() - 12.0 * (1.0 - np.exp(-2.0))) < 0.6


def test_stationary_initialization_replays_and_has_the_stationary_moments():
    # test_scientific_mechanics.jl:133-158 (strict replay), plus the stationary law
    # N(2 s d, 2 s noise) over many cells
    dims = (48, 48)
    n = dims[0] * dims[1]
    component = FluctuatingVolumePressure(eta=1.0, initialization=StationaryMechanicalInitialization, instance_id=7)
    owners = np.arange(1, n + 1, dtype=np.int64).reshape(dims, order="F")
    state = LogicalPottsState(owners, n, cell_types={c: 1 for c in range(1, n + 1)}, medium_domains=(1,),
                              property_schema=required_properties(component))
    state.property_values("target_volume")[:] = 5.0
    state.property_values("volume_strength")[:] = 2.0
    surface = first_shell_relation(SurfaceRole, 2)
    compiled = compile_scientific_state(state, CartesianDomain(dims), BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    spec = build_spec(compiled, first_shell_relation(ProposalRole, 2), ScientificComponentSet(mechanics=(component,)),
                      SequentialCPM(np.float32(3.0)), 0x1234)
    first = bridge.OracleSim(spec, state_arrays(compiled, spec))
    second = bridge.OracleSim(spec, state_arrays(compiled, spec))
    first.initialize_mechanics()
    second.initialize_mechanics()
    q = first.store()["volume_pressure"]
    np.testing.assert_array_equal(q, second.store()["volume_pressure"])
    # displacement 1 - 5 = -4: mean 2 * 2 * -4 = -16, variance 2 * 2 * 3 = 12
    assert abs(q.mean() + 16.0) < 0.25
    assert abs(q.var() - 12.0) < 1.2
    first.step(1)
    second.step(1)
    np.testing.assert_array_equal(first.store()["volume_pressure"], second.store()["volume_pressure"])


def test_mechanics_validation():
    relation = first_shell_relation(SurfaceRole, 2)
    with pytest.raises(ValueError):
        FluctuatingVolumePressure(eta=0.0)
    with pytest.raises(ValueError):
        FluctuatingSurfaceTension(BoundaryEdgeCount(), relation, instance_id=0)
    with pytest.raises(ValueError):
        FixedMechanicalNoise(-1.0)
    volume, tension, surface = _mechanics_pair()
    compiled, proposal, _ = _clock_fixture(SequentialCPM(np.float32(1.0)))
    dup = FluctuatingSurfaceTension(BoundaryEdgeCount(), surface, noise=FixedMechanicalNoise(0.0),
                                    initialization=PreserveMechanicalInitialization, instance_id=1)
    with pytest.raises(ValueError, match="instance"):
        build_spec(compiled, proposal, ScientificComponentSet(mechanics=(volume, dup)), SequentialCPM(np.float32(1.0)))
    dup_state = FluctuatingVolumePressure(noise=FixedMechanicalNoise(0.0), instance_id=3)
    with pytest.raises(ValueError, match="state"):
        build_spec(compiled, proposal, ScientificComponentSet(mechanics=(volume, dup_state)),
                   SequentialCPM(np.float32(1.0)))
    weighted = first_shell_relation(SurfaceRole, 2, weights=(2.0,) * 4)
    mismatched = FluctuatingSurfaceTension(BoundaryEdgeCount(), weighted, noise=FixedMechanicalNoise(0.0), instance_id=4)
    with pytest.raises(ValueError, match="tracker"):
        build_spec(compiled, proposal, ScientificComponentSet(mechanics=(mismatched,)), SequentialCPM(np.float32(1.0)))
