"""CanonicalCheckpoint: capture, validation, exact restore and files
(lib/CorePotts/src/persistence/scientific.jl:30-55, 84-160, 220-475).

A checkpoint holds the logical state at a completed-MCS boundary (ownership tags/ids,
active flags, generations, cell types, reusable slots, medium ids, property columns), the
MCS counter and the seed. That suffices for an exact continuation: the semantic RNG is a
pure function of (seed, mcs, address), so a restored integrator reproduces the original
trajectory bit for bit.

Digests follow the reference's canonical encoding (`_canonical_write!`, type marker then
value, every marker terminated by 0xFF, arrays in column-major order) hashed with SHA-256.
The struct markers of that encoding are Julia type names, and no stored checkpoint of the
reference fixes a digest, so digest bytes are only compared within this implementation.
"""
from __future__ import annotations

import dataclasses
import hashlib
import json
import os
from typing import NamedTuple, Optional, Tuple

import numpy as np

CHECKPOINT_SCHEMA_VERSION = "1.0.0"     # contracts/versions.jl:26
PERSISTENCE_VERSION = "0.1.0"           # persistence/scientific.jl:2
CELL_OWNER_TAG, MEDIUM_OWNER_TAG = 1, 2  # state/logical.jl:25-26
ZERO_DIGEST = bytes(32)
INDEX_TYPES = ("ownership_tag=UInt8", "ownership_id=UInt32", "active=UInt8",
               "generation=UInt64", "cell_type=UInt32", "slot=UInt32", "mcs=UInt64")


class CheckpointIntegrityError(ValueError):
    pass


class IncompleteCheckpointError(ValueError):
    pass


class CheckpointCompatibilityError(ValueError):
    def __init__(self, field: str, expected: str, observed: str):
        super().__init__(f"checkpoint compatibility mismatch for {field}: expected {expected}, "
                         f"observed {observed}")
        self.field, self.expected, self.observed = field, expected, observed


class Version(str):
    """A VersionNumber value for the canonical encoding."""


class ExactContinuationProfile(NamedTuple):
    backend: str
    algorithm: str
    algorithm_contract: Version
    rng_contract: Version
    runtime_version: Version
    corepotts_version: Version
    numerical_mode: str
    scalar_types: Tuple[str, ...]
    index_types: Tuple[str, ...]
    dependency_identities: Tuple[str, ...]
    backend_runtime_identity: str


class CanonicalPropertyColumn(NamedTuple):
    key: str
    values: np.ndarray


class CanonicalCheckpoint(NamedTuple):
    schema_version: Version
    mcs: int
    phase: int
    dims: Tuple[int, ...]
    capacity: int
    ownership_tags: np.ndarray   # uint8, N sites (first axis fastest)
    ownership_ids: np.ndarray    # uint32
    active: np.ndarray           # uint8[capacity]
    generations: np.ndarray      # uint64[capacity]
    cell_types: np.ndarray       # uint32[capacity]
    reusable_slots: np.ndarray   # uint32[capacity]
    reusable_count: int
    medium_ids: np.ndarray       # uint32
    properties: Tuple[CanonicalPropertyColumn, ...]
    seed: int
    profile: ExactContinuationProfile
    model_fingerprint: bytes
    schema_fingerprint: bytes
    topology_fingerprint: bytes
    initial_state_fingerprint: bytes
    ancestry_fingerprint: bytes
    state_fingerprint: bytes
    warnings: Tuple[str, ...]
    checksum: bytes


# ---------------------------------------------------------------------------
# canonical encoding (persistence/scientific.jl:84-160)
# ---------------------------------------------------------------------------
_JULIA_TYPE = {np.dtype(np.uint8): "UInt8", np.dtype(np.uint16): "UInt16", np.dtype(np.uint32): "UInt32",
               np.dtype(np.uint64): "UInt64", np.dtype(np.int8): "Int8", np.dtype(np.int16): "Int16",
               np.dtype(np.int32): "Int32", np.dtype(np.int64): "Int64", np.dtype(np.float32): "Float32",
               np.dtype(np.float64): "Float64"}
_STRUCT_TYPE = {ExactContinuationProfile: "ExactContinuationProfile",
                CanonicalPropertyColumn: "CanonicalPropertyColumn"}


def _marker(out: list, text: str):
    out.append(text.encode() + b"\xff")


def _float_bits(x) -> str:
    x = np.asarray(x)
    width = x.dtype.itemsize * 8
    bits = int(x.view(np.uint32 if width == 32 else np.uint64))
    return format(bits, f"0{width}b")


def _write_array(out: list, a: np.ndarray):
    _marker(out, "Array")
    _marker(out, "Type")
    tname = _JULIA_TYPE[a.dtype]
    _marker(out, tname)
    _write(out, tuple(np.int64(d) for d in a.shape))
    flat = a.reshape(-1, order="F")
    pre = tname.encode() + b"\xff"
    for start in range(0, flat.size, 1 << 20):
        chunk = flat[start:start + (1 << 20)]
        if a.dtype.kind == "f":
            words = [_float_bits(v).encode() for v in chunk]
        else:
            words = chunk.astype(np.bytes_).tolist()
        if words:
            out.append(pre + (b"\xff" + pre).join(words) + b"\xff")


def _write(out: list, value):
    if value is None:
        _marker(out, "nothing")
    elif isinstance(value, (bool, np.bool_)):
        _marker(out, "true" if value else "false")
    elif isinstance(value, np.ndarray):
        _write_array(out, value)
    elif isinstance(value, np.integer):
        _marker(out, _JULIA_TYPE[value.dtype])
        _marker(out, str(int(value)))
    elif isinstance(value, int):
        _marker(out, "Int64")
        _marker(out, str(value))
    elif isinstance(value, np.floating):
        _marker(out, _JULIA_TYPE[value.dtype])
        _marker(out, _float_bits(value))
    elif isinstance(value, float):
        _write(out, np.float64(value))
    elif isinstance(value, Version):
        _marker(out, "VersionNumber")
        _marker(out, str(value))
    elif isinstance(value, bytes):
        _write(out, tuple(np.uint8(b) for b in value))
    elif isinstance(value, str):
        _marker(out, "String")
        _marker(out, value)
    elif type(value) in _STRUCT_TYPE:
        _marker(out, _STRUCT_TYPE[type(value)])
        for name in value._fields:
            _marker(out, name)
            _write(out, getattr(value, name))
    elif isinstance(value, tuple):
        _marker(out, "Tuple")
        _write(out, len(value))
        for item in value:
            _write(out, item)
    else:
        raise TypeError(f"value of type {type(value).__name__} has no canonical fingerprint encoding")


def canonical_digest(*values) -> bytes:
    """SHA-256 of the canonical encoding of values (persistence/scientific.jl:154-160)."""
    out: list = []
    for v in values:
        _write(out, v)
    h = hashlib.sha256()
    for part in out:
        h.update(part)
    return h.digest()


def _state_digest(tags, ids, active, generations, cell_types, reusable, reusable_count, medium_ids,
                  properties) -> bytes:
    return canonical_digest(tags, ids, active, generations, cell_types, reusable,
                            np.uint32(reusable_count), medium_ids, properties)


def _checkpoint_digest(cp: CanonicalCheckpoint) -> bytes:
    return canonical_digest(*[getattr(cp, f) for f in CanonicalCheckpoint._fields if f != "checksum"])


# ---------------------------------------------------------------------------
# fingerprints of the model this engine runs
# ---------------------------------------------------------------------------
def _relation_parts(rel):
    if rel is None:
        return None
    return (tuple(tuple(int(c) for c in o) for o in rel.offsets), np.asarray(rel.weights, dtype=rel.weight_dtype))


def schema_fingerprint(compiled) -> bytes:
    return canonical_digest(tuple((k, str(v.dtype)) for k, v in compiled.properties.items()))


def topology_fingerprint(spec) -> bytes:
    return canonical_digest(tuple(int(d) for d in spec.dims), tuple(map(tuple, spec.face_kind)),
                            tuple(map(tuple, spec.face_medium)),
                            None if spec.mutable_mask is None else np.asarray(spec.mutable_mask, np.uint8),
                            _relation_parts(spec.proposal), _relation_parts(spec.surface))


def model_fingerprint(spec) -> bytes:
    return canonical_digest(topology_fingerprint(spec), _relation_parts(spec.contact),
                            tuple(int(t) for t in spec.energy_order), bool(spec.volume_f64),
                            bool(spec.contact_f64), bool(spec.boundary_f64), np.asarray(spec.J),
                            tuple(int(o) for o in spec.table_owner), tuple(int(t) for t in spec.table_type),
                            tuple(int(m) for m in spec.medium_ids), np.float32(spec.temperature),
                            int(spec.algorithm), tuple(float(r) for r in spec.retire),
                            tuple((int(k), int(i), int(n), np.float32(e), bool(t), np.float32(z))
                                  for k, i, n, e, t, z in spec.mechanics),
                            *((tuple(int(t) for t in spec.tile_size), int(spec.switching_interval))
                              if spec.algorithm == 3 else ()))


def continuation_profile(spec, compiled) -> ExactContinuationProfile:
    scalar = tuple(sorted({_JULIA_TYPE[np.asarray(v).dtype] for v in compiled.properties.values()}))
    name = {1: "CheckerboardSweepCPM", 2: "LotteryCPM", 3: "TiledCheckerboardCPM"}.get(spec.algorithm,
                                                                                      "SequentialCPM")
    return ExactContinuationProfile("MI355X", name, Version("1.0.0"), Version("1.0.0"), Version("1.0.0"),
                                    Version(PERSISTENCE_VERSION), "native_backend", scalar, INDEX_TYPES,
                                    ("libcpm_mi355x@abi1",), "hip-gfx950")


# ---------------------------------------------------------------------------
# capture / validate / restore (persistence/scientific.jl:250-475)
# ---------------------------------------------------------------------------
def capture_arrays(state: dict, compiled, spec, mcs: int, seed: int,
                   ancestry: Optional[CanonicalCheckpoint] = None) -> CanonicalCheckpoint:
    """Checkpoint of downloaded state arrays (cpm_download_state order) at MCS `mcs`."""
    dims = tuple(int(d) for d in compiled.domain.dims)
    cap = int(compiled.capacity)
    tags = np.ascontiguousarray(state["tags"], np.uint8).reshape(dims, order="F")
    ids = np.ascontiguousarray(state["ids"], np.uint32).reshape(dims, order="F")
    active = np.ascontiguousarray(state["active"], np.uint8)
    generations = np.zeros(cap, np.uint64)   # this path has no lifecycle events
    cell_types = np.ascontiguousarray(state["cell_types"], np.uint32)
    reusable = np.zeros(cap, np.uint32)
    medium_ids = np.asarray(compiled.medium_ids, np.uint32)
    props = tuple(CanonicalPropertyColumn(k, np.asarray(state.get(k, v)).astype(v.dtype))
                  for k, v in compiled.properties.items())
    state_fp = _state_digest(tags, ids, active, generations, cell_types, reusable, 0, medium_ids, props)
    if ancestry is not None:
        validate_checkpoint(ancestry)
        initial = ancestry.initial_state_fingerprint
    elif mcs == 0:
        initial = state_fp
    else:
        initial = ZERO_DIGEST
    warnings = ("initial-state fingerprint unavailable; capture an MCS-0 ancestor for full provenance",) \
        if initial == ZERO_DIGEST else ()
    cp = CanonicalCheckpoint(Version(CHECKPOINT_SCHEMA_VERSION), int(mcs), 1, dims, cap, tags, ids, active,
                             generations, cell_types, reusable, 0, medium_ids, props, int(seed),
                             continuation_profile(spec, compiled), model_fingerprint(spec),
                             schema_fingerprint(compiled), topology_fingerprint(spec), initial,
                             ZERO_DIGEST if ancestry is None else ancestry.checksum, state_fp, warnings,
                             ZERO_DIGEST)
    return cp._replace(checksum=_checkpoint_digest(cp))


def capture_checkpoint(integrator, ancestry: Optional[CanonicalCheckpoint] = None) -> CanonicalCheckpoint:
    """capture_checkpoint(integrator; ancestry) (persistence/scientific.jl:250-292)."""
    return capture_arrays(integrator.download(), integrator.compiled, integrator.spec, integrator.mcs,
                          integrator.seed, ancestry)


def validate_checkpoint(cp: CanonicalCheckpoint) -> CanonicalCheckpoint:
    """validate_checkpoint (persistence/scientific.jl:294-365)."""
    if cp.schema_version != CHECKPOINT_SCHEMA_VERSION:
        raise CheckpointCompatibilityError("schema_version", CHECKPOINT_SCHEMA_VERSION, str(cp.schema_version))
    if cp.phase != 1:
        raise CheckpointIntegrityError("checkpoint is not at a completed-MCS boundary")
    if not all(d > 0 for d in cp.dims):
        raise CheckpointIntegrityError("checkpoint dimensions must be positive")
    if int(np.prod(cp.dims)) != cp.ownership_ids.size:
        raise CheckpointIntegrityError("checkpoint ownership dimensions are inconsistent")
    if cp.ownership_tags.size != cp.ownership_ids.size:
        raise CheckpointIntegrityError("checkpoint ownership tag/id lengths differ")
    cap = int(cp.capacity)
    if not all(len(x) == cap for x in (cp.active, cp.generations, cp.cell_types, cp.reusable_slots)):
        raise CheckpointIntegrityError("checkpoint fixed-capacity columns have inconsistent lengths")
    if cp.reusable_count > cp.capacity:
        raise CheckpointIntegrityError("checkpoint reusable count exceeds capacity")
    if not np.all((cp.active == 0) | (cp.active == 1)):
        raise CheckpointIntegrityError("checkpoint active flags must be zero or one")
    if cp.medium_ids.size == 0:
        raise CheckpointIntegrityError("checkpoint must declare at least one medium identity")
    if np.unique(cp.medium_ids).size != cp.medium_ids.size:
        raise CheckpointIntegrityError("checkpoint medium identities must be unique")
    if np.any(cp.medium_ids == 0):
        raise CheckpointIntegrityError("checkpoint medium identities must be positive")
    reusable = cp.reusable_slots[:cp.reusable_count]
    if np.unique(reusable).size != reusable.size:
        raise CheckpointIntegrityError("checkpoint reusable identities must be unique")
    if not all(0 < s <= cap and cp.active[int(s) - 1] == 0 for s in reusable):
        raise CheckpointIntegrityError("checkpoint reusable identities must name inactive slots within capacity")
    if np.any(cp.reusable_slots[cp.reusable_count:] != 0):
        raise CheckpointIntegrityError("checkpoint reusable padding must be canonical zero storage")
    tags = cp.ownership_tags.reshape(-1, order="F")
    ids = cp.ownership_ids.reshape(-1, order="F").astype(np.int64)
    cells = tags == CELL_OWNER_TAG
    media = tags == MEDIUM_OWNER_TAG
    if np.any(~(cells | media)):
        raise CheckpointIntegrityError("checkpoint contains an invalid ownership tag")
    cid = ids[cells]
    if cid.size and (np.any(cid < 1) or np.any(cid > cap) or np.any(cp.active[cid - 1] == 0)):
        raise CheckpointIntegrityError("checkpoint ownership names an inactive or out-of-capacity cell")
    if not np.all(np.isin(ids[media], cp.medium_ids.astype(np.int64))):
        raise CheckpointIntegrityError("checkpoint ownership names an undeclared medium identity")
    keys = [c.key for c in cp.properties]
    if len(set(keys)) != len(keys):
        raise CheckpointIntegrityError("checkpoint property keys must be unique")
    if not all(len(c.values) == cap for c in cp.properties):
        raise CheckpointIntegrityError("checkpoint property columns must match fixed cell capacity")
    expected = _state_digest(cp.ownership_tags, cp.ownership_ids, cp.active, cp.generations, cp.cell_types,
                             cp.reusable_slots, cp.reusable_count, cp.medium_ids, cp.properties)
    if expected != cp.state_fingerprint:
        raise CheckpointIntegrityError("checkpoint state fingerprint does not match authoritative state")
    if _checkpoint_digest(cp) != cp.checksum:
        raise CheckpointIntegrityError("checkpoint canonical checksum does not match its payload")
    return cp


def checkpoint_compiled_state(cp: CanonicalCheckpoint, prototype_compiled):
    """The compiled state a checkpoint describes: ownership, active cells and their types,
    property columns, trackers rebuilt from ownership and obstacle owners checked
    (checkpoint_logical_state + compile_scientific_state, persistence/scientific.jl:359-397, 430-433)."""
    from .model import LogicalPottsState, compile_scientific_state, owner_words
    validate_checkpoint(cp)
    if schema_fingerprint(prototype_compiled) != cp.schema_fingerprint:
        raise CheckpointCompatibilityError("property_schema", cp.schema_fingerprint.hex(),
                                           schema_fingerprint(prototype_compiled).hex())
    if tuple(c.key for c in cp.properties) != tuple(prototype_compiled.properties):
        raise CheckpointIntegrityError("checkpoint property keys do not match the compatible schema")
    tags = cp.ownership_tags.reshape(-1, order="F")
    ids = cp.ownership_ids.reshape(-1, order="F")
    active_ids = np.flatnonzero(cp.active) + 1
    logical = LogicalPottsState(owner_words(tags, ids), cp.capacity,
                                cell_types={int(c): int(cp.cell_types[c - 1]) for c in active_ids},
                                medium_domains=tuple(int(m) for m in cp.medium_ids),
                                property_schema={k: v.dtype for k, v in prototype_compiled.properties.items()},
                                dims=cp.dims, active_ids=active_ids)
    for c in cp.properties:
        want = prototype_compiled.properties[c.key].dtype
        if c.values.dtype != want:
            raise CheckpointCompatibilityError(c.key, str(want), str(c.values.dtype))
        logical.properties[c.key][:] = c.values
    return compile_scientific_state(logical, prototype_compiled.domain, prototype_compiled.tracker)


def _check_prototype(cp: CanonicalCheckpoint, prototype, exact: bool):
    spec = prototype.spec
    if cp.topology_fingerprint != topology_fingerprint(spec):
        raise CheckpointCompatibilityError("topology", "prototype topology", "checkpoint topology")
    if cp.model_fingerprint != model_fingerprint(spec):
        raise CheckpointCompatibilityError("model", "prototype model fingerprint", "checkpoint model fingerprint")
    expected = continuation_profile(spec, prototype.compiled)
    if exact and cp.profile != expected:
        raise CheckpointCompatibilityError("continuation_profile", str(expected), str(cp.profile))


def _restore(cp: CanonicalCheckpoint, prototype, exact: bool):
    from .native import ScientificPottsIntegrator
    validate_checkpoint(cp)
    if schema_fingerprint(prototype.compiled) != cp.schema_fingerprint:
        raise CheckpointCompatibilityError("property_schema", cp.schema_fingerprint.hex(),
                                           schema_fingerprint(prototype.compiled).hex())
    _check_prototype(cp, prototype, exact)
    compiled = checkpoint_compiled_state(cp, prototype.compiled)
    integ = ScientificPottsIntegrator(compiled, dataclasses.replace(prototype.spec, seed=int(cp.seed)))
    integ.mcs = cp.mcs
    return integ


def restore_checkpoint(cp: CanonicalCheckpoint, prototype):
    """restore_checkpoint(checkpoint, prototype) (persistence/scientific.jl:464-468): a new
    integrator of the prototype's model at the checkpoint's state, MCS and seed, with the
    exact-continuation profile required to match."""
    return _restore(cp, prototype, exact=True)


class LogicalImportReport(NamedTuple):
    checksum: bytes
    source_backend: str
    target_backend: str
    guarantee: str
    warnings: Tuple[str, ...]


def import_checkpoint(cp: CanonicalCheckpoint, prototype):
    """import_checkpoint (persistence/scientific.jl:470-477): restore without the profile
    check; the report states that exact continuation is not claimed."""
    integ = _restore(cp, prototype, exact=False)
    return integ, LogicalImportReport(cp.checksum, cp.profile.backend, "MI355X", "statistical_or_weaker",
                                      ("logical import does not claim exact continuation",))


def restore_into_batch(cp: CanonicalCheckpoint, batch, spec, prototype_compiled,
                       target_mcs: Optional[int] = None) -> int:
    """One more replica of `batch` at the checkpoint's state, seed and MCS; returns its slot."""
    if cp.topology_fingerprint != topology_fingerprint(spec) or cp.model_fingerprint != model_fingerprint(spec):
        raise CheckpointCompatibilityError("model", "replica model fingerprint", "checkpoint model fingerprint")
    compiled = checkpoint_compiled_state(cp, prototype_compiled)
    slot = batch.reserve(dataclasses.replace(spec, seed=int(cp.seed)))
    batch.load(slot, compiled, initialize_mechanics=False)
    batch.set_mcs(slot, cp.mcs)
    if target_mcs is not None:
        batch.set_target(slot, target_mcs)
    return slot


def capture_batch_checkpoint(batch, slot: int, ancestry: Optional[CanonicalCheckpoint] = None):
    """Checkpoint of one batch replica at its current MCS (synchronizes)."""
    spec = batch.specs[slot]
    mcs = batch.progress()[slot]
    return capture_arrays(batch.download(slot), batch.compiled[slot], spec, mcs, spec.seed, ancestry)


# ---------------------------------------------------------------------------
# files: numpy .npz, no pickled objects
# ---------------------------------------------------------------------------
def save_checkpoint(cp: CanonicalCheckpoint, path: str):
    """Write a checkpoint file. The payload goes to a temporary name first and replaces `path`
    only when complete, so a failed write leaves any earlier checkpoint at `path` intact
    (the all-or-nothing contract of write_checkpoint!, persistence/scientific.jl:490-500)."""
    validate_checkpoint(cp)
    meta = {"complete": 1, "schema_version": str(cp.schema_version), "mcs": cp.mcs, "phase": cp.phase, "dims": list(cp.dims),
            "capacity": cp.capacity, "reusable_count": cp.reusable_count, "seed": str(cp.seed),
            "profile": [list(v) if isinstance(v, tuple) else str(v) for v in cp.profile],
            "property_keys": [c.key for c in cp.properties], "warnings": list(cp.warnings)}
    for f in ("model_fingerprint", "schema_fingerprint", "topology_fingerprint", "initial_state_fingerprint",
              "ancestry_fingerprint", "state_fingerprint", "checksum"):
        meta[f] = getattr(cp, f).hex()
    arrays = {"ownership_tags": cp.ownership_tags, "ownership_ids": cp.ownership_ids, "active": cp.active,
              "generations": cp.generations, "cell_types": cp.cell_types, "reusable_slots": cp.reusable_slots,
              "medium_ids": cp.medium_ids}
    for i, c in enumerate(cp.properties):
        arrays[f"property_{i}"] = c.values
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    partial = path + ".partial"
    try:
        with open(partial, "wb") as handle:
            np.savez(handle, **arrays)
            handle.flush()
            os.fsync(handle.fileno())
        os.replace(partial, path)
    finally:
        if os.path.exists(partial):
            os.remove(partial)


def load_checkpoint(path: str) -> CanonicalCheckpoint:
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(bytes(z["meta"]).decode())
        if meta.get("complete") != 1:
            raise IncompleteCheckpointError(f"checkpoint file {path} is incomplete")
        p = meta["profile"]
        profile = ExactContinuationProfile(p[0], p[1], Version(p[2]), Version(p[3]), Version(p[4]), Version(p[5]),
                                           p[6], tuple(p[7]), tuple(p[8]), tuple(p[9]), p[10])
        props = tuple(CanonicalPropertyColumn(k, z[f"property_{i}"]) for i, k in enumerate(meta["property_keys"]))
        cp = CanonicalCheckpoint(Version(meta["schema_version"]), int(meta["mcs"]), int(meta["phase"]),
                                 tuple(meta["dims"]), int(meta["capacity"]), z["ownership_tags"], z["ownership_ids"],
                                 z["active"], z["generations"], z["cell_types"], z["reusable_slots"],
                                 int(meta["reusable_count"]), z["medium_ids"], props, int(meta["seed"]), profile,
                                 *[bytes.fromhex(meta[f]) for f in ("model_fingerprint", "schema_fingerprint",
                                                                    "topology_fingerprint", "initial_state_fingerprint",
                                                                    "ancestry_fingerprint", "state_fingerprint")],
                                 tuple(meta["warnings"]), bytes.fromhex(meta["checksum"]))
    return validate_checkpoint(cp)
