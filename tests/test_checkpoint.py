"""CanonicalCheckpoint on the CPU, following lib/CorePotts/test/test_phase8_persistence.jl:
capture at MCS 0 and at a later boundary with ancestry, validation of every corrupted field the
reference tests, compatibility errors, file round trips and exact continuation.

Exact continuation is checked with the oracle as the integrator: a run interrupted at MCS 2,
captured, restored through the checkpoint (trackers rebuilt from ownership) and continued for 3
MCS equals the uninterrupted 5-MCS run in lattice, trackers, property columns and report.
Digest bytes have no reference-held value to compare with (parity unpinned); the tests pin the
encoding's structure and the digests' sensitivity instead.
"""
import dataclasses
import types

import numpy as np
import pytest

from corepotts import (BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain, CellOwner,
                       CheckerboardSweepCPM, ContactRole, LogicalPottsState, LotteryCPM, MediumOwner,
                       MediumTypeTable, ProposalRole, QuadraticBoundaryHamiltonian,
                       QuadraticVolumeHamiltonian, ScientificComponentSet, SurfaceRole,
                       UnorderedContactHamiltonian, build_spec, compile_scientific_state,
                       first_shell_relation, moore_relation, required_properties, state_arrays)
from corepotts import checkpoint as ck
from oracle import bridge

SEED = 0x7065727369737401


def phase8_model(algorithm):
    """test_phase8_persistence.jl:8-30 without its lifecycle event (lifecycle is not on this path)."""
    energy = QuadraticVolumeHamiltonian(number_type=np.float32)
    owners = np.full((6, 6), MediumOwner(1), dtype=np.int64)
    owners[1:5, 1:3] = CellOwner(1)
    owners[1:5, 3:5] = CellOwner(2)
    logical = LogicalPottsState(owners, 4, cell_types={1: 1, 2: 2}, medium_domains=(1,),
                                property_schema=required_properties(energy))
    logical.property_values("target_volume")[:2] = np.float32(8)
    logical.property_values("volume_strength")[:2] = np.float32(1)
    tracker = BoundaryMeasureTracker(BoundaryEdgeCount(), first_shell_relation(SurfaceRole, 2))
    compiled = compile_scientific_state(logical, CartesianDomain((6, 6)), tracker)
    components = ScientificComponentSet(energies=(energy,))
    spec = build_spec(compiled, first_shell_relation(ProposalRole, 2), components, algorithm, SEED)
    return compiled, spec


def tissue_model(algorithm, seed=0x5EED):
    """A 24x20 tissue of 12 cells with volume, contact (Moore) and boundary energies."""
    dims = (24, 20)
    owners = np.full(dims, MediumOwner(1), dtype=np.int64)
    cid = 0
    for bx in range(4):
        for by in range(3):
            cid += 1
            owners[2 + 5 * bx:6 + 5 * bx, 2 + 6 * by:7 + 6 * by] = CellOwner(cid)
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    surface = first_shell_relation(SurfaceRole, 2)
    boundary = QuadraticBoundaryHamiltonian(BoundaryEdgeCount(), surface)
    J = np.array([[0.0, 8.0, 12.0], [8.0, 4.0, 10.0], [12.0, 10.0, 6.0]])
    contact = UnorderedContactHamiltonian(J, MediumTypeTable({1: 1}), moore_relation(ContactRole, 2))
    schema = {**required_properties(volume), **required_properties(boundary)}
    logical = LogicalPottsState(owners, 16, cell_types={c: 2 + c % 2 for c in range(1, 13)},
                                medium_domains=(1,), property_schema=schema)
    logical.property_values("target_volume")[:12] = np.float32(22)
    logical.property_values("volume_strength")[:12] = np.float32(1.5)
    logical.property_values("target_boundary")[:12] = 18
    logical.property_values("boundary_strength")[:12] = 0.25
    compiled = compile_scientific_state(logical, CartesianDomain(dims), BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    components = ScientificComponentSet(energies=(volume, contact, boundary))
    spec = build_spec(compiled, first_shell_relation(ProposalRole, 2), components, algorithm, seed)
    return compiled, spec


def oracle_capture(sim, compiled, spec, ancestry=None):
    return ck.capture_arrays(sim.store(), compiled, spec, sim.mcs, spec.seed, ancestry)


def oracle_restore(cp, compiled, spec):
    restored = ck.checkpoint_compiled_state(cp, compiled)
    rspec = dataclasses.replace(spec, seed=cp.seed)
    sim = bridge.OracleSim(rspec, state_arrays(restored, rspec))
    sim.mcs = cp.mcs
    return sim, restored


ALGORITHMS = [CheckerboardSweepCPM(np.float32(2.0)), LotteryCPM(np.float32(2.0))]


def test_root_checkpoint_fields():
    compiled, spec = phase8_model(ALGORITHMS[0])
    sim = bridge.OracleSim(spec, state_arrays(compiled, spec))
    root = oracle_capture(sim, compiled, spec)
    assert root.mcs == 0 and root.phase == 1
    assert root.initial_state_fingerprint != ck.ZERO_DIGEST
    assert root.initial_state_fingerprint == root.state_fingerprint
    assert root.schema_version == ck.CHECKPOINT_SCHEMA_VERSION
    assert root.profile.numerical_mode == "native_backend"
    assert "Float32" in root.profile.scalar_types
    assert "ownership_id=UInt32" in root.profile.index_types
    assert root.warnings == ()
    assert ck.validate_checkpoint(root) is root
    assert root.ownership_tags.shape == (6, 6)
    assert root.ownership_ids[1, 1] == 1 and root.ownership_ids[1, 3] == 2
    np.testing.assert_array_equal(root.active, [1, 1, 0, 0])
    np.testing.assert_array_equal(root.cell_types, [1, 2, 0, 0])
    assert [c.key for c in root.properties] == ["target_volume", "volume_strength"]
    assert root.properties[0].values.dtype == np.float32


def _corrupt(cp, **fields):
    return cp._replace(**fields)


def test_validation_rejects_corruption():
    compiled, spec = phase8_model(ALGORITHMS[0])
    root = oracle_capture(bridge.OracleSim(spec, state_arrays(compiled, spec)), compiled, spec)
    active = root.active.copy()
    active[0] ^= 1
    with pytest.raises(ck.CheckpointIntegrityError):
        ck.validate_checkpoint(_corrupt(root, active=active))
    with pytest.raises(ck.CheckpointCompatibilityError):
        ck.validate_checkpoint(_corrupt(root, schema_version=ck.Version("999.0.0")))
    with pytest.raises(ck.CheckpointIntegrityError):
        ck.validate_checkpoint(_corrupt(root, capacity=root.capacity + 1))
    tags = root.ownership_tags.copy()
    tags[0, 0] = 99
    with pytest.raises(ck.CheckpointIntegrityError, match="invalid ownership tag"):
        ck.validate_checkpoint(_corrupt(root, ownership_tags=tags))
    props = (ck.CanonicalPropertyColumn("target_volume", root.properties[0].values[:-1]), root.properties[1])
    with pytest.raises(ck.CheckpointIntegrityError, match="fixed cell capacity"):
        ck.validate_checkpoint(_corrupt(root, properties=props))
    ids = root.ownership_ids.copy()
    ids[1, 1] = 3                                   # names an inactive cell
    with pytest.raises(ck.CheckpointIntegrityError, match="inactive"):
        ck.validate_checkpoint(_corrupt(root, ownership_ids=ids))
    with pytest.raises(ck.CheckpointIntegrityError, match="completed-MCS"):
        ck.validate_checkpoint(_corrupt(root, phase=2))
    with pytest.raises(ck.CheckpointIntegrityError, match="reusable padding"):
        ck.validate_checkpoint(_corrupt(root, reusable_slots=np.array([0, 0, 3, 0], np.uint32)))
    # a consistent state with a stale checksum
    vols = root.properties[0].values.copy()
    vols[0] = 9
    p2 = (ck.CanonicalPropertyColumn("target_volume", vols), root.properties[1])
    with pytest.raises(ck.CheckpointIntegrityError, match="state fingerprint"):
        ck.validate_checkpoint(_corrupt(root, properties=p2))
    with pytest.raises(ck.CheckpointIntegrityError, match="checksum"):
        ck.validate_checkpoint(_corrupt(root, mcs=7))


def test_property_schema_mismatch():
    compiled, spec = phase8_model(ALGORITHMS[0])
    root = oracle_capture(bridge.OracleSim(spec, state_arrays(compiled, spec)), compiled, spec)
    other = compiled.copy()
    other.properties = {"other_target": other.properties["target_volume"],
                        "volume_strength": other.properties["volume_strength"]}
    with pytest.raises(ck.CheckpointCompatibilityError):
        ck.checkpoint_compiled_state(root, other)


@pytest.mark.parametrize("model", [phase8_model, tissue_model])
@pytest.mark.parametrize("algorithm", ALGORITHMS, ids=["checkerboard", "lottery"])
def test_exact_continuation_through_checkpoint(model, algorithm, tmp_path):
    compiled, spec = model(algorithm)
    uninterrupted = bridge.OracleSim(spec, state_arrays(compiled, spec))
    root = oracle_capture(uninterrupted, compiled, spec)
    uninterrupted.step(2)
    cp = oracle_capture(uninterrupted, compiled, spec, ancestry=root)
    assert cp.initial_state_fingerprint == root.initial_state_fingerprint
    assert cp.ancestry_fingerprint == root.checksum
    assert cp.state_fingerprint != root.state_fingerprint
    uninterrupted.step(3)
    expected = uninterrupted.store()

    path = str(tmp_path / "cp.npz")
    ck.save_checkpoint(cp, path)
    loaded = ck.load_checkpoint(path)
    assert loaded.checksum == cp.checksum
    resumed, restored = oracle_restore(loaded, compiled, spec)
    # trackers rebuilt from ownership equal the tracked values at the interruption
    at_cp = bridge.OracleSim(spec, state_arrays(compiled, spec))
    at_cp.step(2)
    mid = at_cp.store()
    np.testing.assert_array_equal(restored.finite_volumes, mid["finite_volumes"])
    np.testing.assert_array_equal(restored.boundary_measures, mid["boundary_measures"])
    np.testing.assert_array_equal(restored.medium_volumes, mid["medium_volumes"])
    assert resumed.mcs == 2
    resumed.step(3)
    observed = resumed.store()
    assert resumed.mcs == uninterrupted.mcs == 5
    for key in expected:
        np.testing.assert_array_equal(observed[key], expected[key], err_msg=key)
    assert resumed.report() == uninterrupted.report()


def test_restore_requires_same_model_and_import_does_not():
    compiled, spec = phase8_model(ALGORITHMS[0])
    sim = bridge.OracleSim(spec, state_arrays(compiled, spec))
    sim.step(1)
    cp = oracle_capture(sim, compiled, spec)
    assert cp.warnings                             # no MCS-0 ancestor
    _, lottery_spec = phase8_model(ALGORITHMS[1])
    different = types.SimpleNamespace(spec=lottery_spec, compiled=compiled)
    with pytest.raises(ck.CheckpointCompatibilityError):
        ck.restore_checkpoint(cp, different)
    hotter = dataclasses.replace(spec, temperature=np.float32(3.0))
    with pytest.raises(ck.CheckpointCompatibilityError, match="model"):
        ck.restore_checkpoint(cp, types.SimpleNamespace(spec=hotter, compiled=compiled))
    # same model: the continuation profile matches
    ck._check_prototype(cp, types.SimpleNamespace(spec=spec, compiled=compiled), exact=True)


def test_incomplete_and_failed_files(tmp_path):
    compiled, spec = phase8_model(ALGORITHMS[0])
    root = oracle_capture(bridge.OracleSim(spec, state_arrays(compiled, spec)), compiled, spec)
    path = str(tmp_path / "root.npz")
    ck.save_checkpoint(root, path)
    with np.load(path) as z:
        arrays = {k: z[k] for k in z.files}
    import json
    meta = json.loads(bytes(arrays["meta"]).decode())
    meta["complete"] = 0
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), np.uint8)
    bad = str(tmp_path / "bad.npz")
    np.savez(bad, **arrays)
    with pytest.raises(ck.IncompleteCheckpointError):
        ck.load_checkpoint(bad)
    # an invalid checkpoint is never written and the earlier file survives
    with pytest.raises(ck.CheckpointIntegrityError):
        ck.save_checkpoint(root._replace(mcs=3), path)
    assert ck.load_checkpoint(path).checksum == root.checksum


def test_canonical_encoding_structure():
    # markers are UTF-8 text followed by 0xFF; integers carry their type, floats their bit string
    out = []
    ck._write(out, 5)
    assert b"".join(out) == b"Int64\xff5\xff"
    out = []
    ck._write(out, np.float32(1.0))
    assert b"".join(out) == b"Float32\xff" + b"00111111100000000000000000000000\xff"
    out = []
    ck._write(out, np.array([[1, 2], [3, 4]], np.uint8))   # column-major items
    enc = b"".join(out)
    assert enc.startswith(b"Array\xffType\xffUInt8\xffTuple\xffInt64\xff2\xffInt64\xff2\xffInt64\xff2\xff")
    assert enc.endswith(b"UInt8\xff1\xffUInt8\xff3\xffUInt8\xff2\xffUInt8\xff4\xff")
    # digests see element order, element type and shape
    a = np.arange(6, dtype=np.uint32)
    d = ck.canonical_digest(a)
    assert d != ck.canonical_digest(a[::-1].copy())
    assert d != ck.canonical_digest(a.astype(np.uint64))
    assert d != ck.canonical_digest(a.reshape(2, 3))
    assert len(d) == 32
