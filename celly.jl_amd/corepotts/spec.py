"""Lowering of a compiled state + components + algorithm into the flat model description
that the C-ABI engine (and, in tests, the CPU oracle) consume.

Mirrors the validation init_scientific performs before allocating a workspace
(lib/CorePotts/src/algorithms/sequential.jl:389-416, checkerboard.jl:136-148,
lottery.jl:30-40) and the arithmetic-type promotion of each energy term
(scientific_trackers.jl:123-172, scientific_components.jl:289-308).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from .model import (MECH_PRESSURE, MECH_STATE_COLUMN, MECH_TENSION, BoundaryEdgeCount,
                    CheckerboardSweepCPM, ClosedBoundary, CompiledScientificState, FixedExterior,
                    FluctuatingSurfaceTension, FluctuatingVolumePressure, LotteryCPM, PeriodicBoundary,
                    TiledCheckerboardCPM,
                    QuadraticBoundaryHamiltonian, QuadraticVolumeHamiltonian, SequentialCPM,
                    StaticCartesianRelation, UnorderedContactHamiltonian,
                    validate_relation_domain)

TERM_VOLUME, TERM_CONTACT, TERM_BOUNDARY = 0, 1, 2
ALG_SEQUENTIAL, ALG_CHECKERBOARD, ALG_LOTTERY, ALG_TILED = 0, 1, 2, 3
MAX_TYPES = 16
MAX_MEDIA = 16


@dataclass
class ModelSpec:
    algorithm: int
    ndim: int
    dims: Tuple[int, int, int]
    face_kind: List[List[int]]
    face_medium: List[List[int]]
    mutable_mask: Optional[np.ndarray]
    immutable_tags: Optional[np.ndarray]
    immutable_ids: Optional[np.ndarray]
    proposal: StaticCartesianRelation
    contact: Optional[StaticCartesianRelation]
    surface: StaticCartesianRelation
    energy_order: List[int]
    volume_f64: bool
    contact_f64: bool
    boundary_f64: bool
    J: np.ndarray
    table_owner: Tuple[int, ...]
    table_type: Tuple[int, ...]
    medium_ids: Tuple[int, ...]
    temperature: np.float32
    capacity: int
    seed: int
    retire: Tuple[float, float, int, float] = (0.0, 0.0, 0, 0.0)
    columns: dict = field(default_factory=dict)   # volume/boundary column -> property key
    # (kind, initialization, instance_id, eta, noise_is_temperature, noise) per mechanical
    # component, ScientificComponentSet.mechanics order
    mechanics: tuple = ()
    tile_size: Tuple[int, int, int] = (0, 0, 0)   # TiledCheckerboardCPM (0: default extents)
    switching_interval: int = 0                   # TiledCheckerboardCPM (0: none)

    @property
    def n_types(self) -> int:
        return int(self.J.shape[0]) if self.J.size else 0


def _face(boundary) -> Tuple[int, int]:
    if isinstance(boundary, PeriodicBoundary):
        return 0, 0
    if isinstance(boundary, ClosedBoundary):
        return 1, 0
    if isinstance(boundary, FixedExterior):
        return 2, -boundary.owner
    raise TypeError("unknown boundary condition")


def algorithm_id(algorithm) -> int:
    if isinstance(algorithm, CheckerboardSweepCPM):
        return ALG_CHECKERBOARD
    if isinstance(algorithm, LotteryCPM):
        return ALG_LOTTERY
    if isinstance(algorithm, SequentialCPM):
        return ALG_SEQUENTIAL
    if isinstance(algorithm, TiledCheckerboardCPM):
        return ALG_TILED
    raise TypeError("unsupported algorithm")


def _mechanics_entry(component, compiled, algorithm, columns):
    """(kind, initialization, instance_id, eta, noise_is_temperature, noise) of a mechanical
    component (scientific_components.jl:152-182, 576-616), after the checks init_scientific
    makes on it; its target and strength columns join the energies' columns."""
    props = compiled.properties
    if isinstance(component, FluctuatingVolumePressure):
        kind, target_name, strength_name = MECH_PRESSURE, "target_volume", "volume_strength"
    elif isinstance(component, FluctuatingSurfaceTension):
        if not isinstance(component.metric, BoundaryEdgeCount):
            raise NotImplementedError("only BoundaryEdgeCount surface tension runs on this engine")
        if component.relation != compiled.tracker.relation:
            raise ValueError("surface tension does not match the compiled boundary tracker descriptor")
        kind, target_name, strength_name = MECH_TENSION, "target_boundary", "boundary_strength"
    else:
        raise NotImplementedError(f"mechanics {type(component).__name__} is outside this engine")
    if np.dtype(component.number_type) != np.float32:
        raise NotImplementedError("the engine keeps mechanical state in Float32")
    if component.state != MECH_STATE_COLUMN[kind]:
        raise NotImplementedError(f"the engine keeps this mechanical state in the {MECH_STATE_COLUMN[kind]} column")
    for name, key in ((target_name, component.target), (strength_name, component.strength)):
        if columns.setdefault(name, key) != key:
            raise NotImplementedError(f"mechanics and energies must share the {name} column")
        if key not in props:
            raise ValueError(f"state lacks the property column {key}")
    if component.state not in props:
        raise ValueError(f"state lacks the property column {component.state}")
    noise_is_temperature = component.noise is None
    noise = np.float32(0.0) if noise_is_temperature else np.float32(component.noise.value)
    return (kind, component.initialization, int(component.instance_id), np.float32(component.eta),
            noise_is_temperature, noise)


def julia_float_promotion(*dtypes) -> np.dtype:
    """float(promote_type(T...)) under Julia's rules, the arithmetic type of an energy term
    (scientific_components.jl:101, 112, 529, 549): an integer column promotes with a float one
    to that float type (Int64 with Float32 gives Float32, where numpy would give float64),
    Float32 with Float64 gives Float64, and integers alone give Float64."""
    floats = [np.dtype(d) for d in dtypes if np.issubdtype(np.dtype(d), np.floating)]
    if not floats:
        return np.dtype(np.float64)
    return np.dtype(np.float64) if any(d.itemsize > 4 for d in floats) else np.dtype(np.float32)


def build_spec(compiled: CompiledScientificState, proposal_relation: StaticCartesianRelation,
               components, algorithm, seed: int = 0) -> ModelSpec:
    if not 0 <= int(seed) <= 0xFFFFFFFFFFFFFFFF:
        raise ValueError("scientific integrator seed must fit UInt64")
    if proposal_relation.role.name != "ProposalRole":
        raise ValueError("the proposal relation must carry ProposalRole")
    alg = algorithm_id(algorithm)
    name = type(algorithm).__name__
    if alg != ALG_SEQUENTIAL and components.constraints:
        raise ValueError(f"{name} does not yet qualify hard constraints; use SequentialCPM")
    if components.drives or components.kinetic_modifiers:
        raise NotImplementedError("drives and kinetic modifiers are outside this engine")
    if components.constraints:
        raise NotImplementedError("hard constraints are outside this engine")
    if np.dtype(type(algorithm.temperature)) != np.float32:
        raise NotImplementedError("the engine evaluates acceptance with a Float32 temperature")
    domain = compiled.domain
    validate_relation_domain(domain, proposal_relation)
    mutable_count = domain.mutable_site_count()
    if not 0 < mutable_count <= 0xFFFFFFFF:
        raise ValueError("scientific semantic addressing requires 1:typemax(UInt32) mutable sites")

    order: List[int] = []
    volume_f64 = contact_f64 = boundary_f64 = False
    contact_rel = None
    J = np.zeros((0, 0), dtype=np.float64)
    table_owner: Tuple[int, ...] = ()
    table_type: Tuple[int, ...] = ()
    columns = {}
    props = compiled.properties
    for component in components.energies:
        if isinstance(component, QuadraticVolumeHamiltonian):
            order.append(TERM_VOLUME)
            tv, vs = props[component.target], props[component.strength]
            volume_f64 = julia_float_promotion(tv.dtype, vs.dtype) == np.float64
            columns["target_volume"] = component.target
            columns["volume_strength"] = component.strength
        elif isinstance(component, UnorderedContactHamiltonian):
            order.append(TERM_CONTACT)
            contact_rel = component.relation
            validate_relation_domain(domain, contact_rel)
            J = np.asarray(component.interactions)
            contact_f64 = J.dtype == np.float64
            table_owner = component.medium_types.owner_ids
            table_type = component.medium_types.type_ids
        elif isinstance(component, QuadraticBoundaryHamiltonian):
            if not isinstance(component.metric, BoundaryEdgeCount):
                raise NotImplementedError("only BoundaryEdgeCount boundary energies run on this engine")
            if component.relation != compiled.tracker.relation:
                raise ValueError("boundary component does not match the compiled boundary tracker descriptor")
            order.append(TERM_BOUNDARY)
            tb, bs = props[component.target], props[component.strength]
            boundary_f64 = julia_float_promotion(tb.dtype, bs.dtype, component.relation.weight_dtype) == np.float64
            columns["target_boundary"] = component.target
            columns["boundary_strength"] = component.strength
        else:
            raise NotImplementedError(f"energy {type(component).__name__} is outside this engine")
    if len(set(order)) != len(order):
        raise NotImplementedError("each energy term may appear once")
    tile_size, switching = (0, 0, 0), 0
    if alg == ALG_TILED:
        # tiled_checkerboard_device.jl:132-163
        if components.mechanics or components.constraints:
            raise ValueError("the resident tiled path rejects constraints and mechanical components")
        if algorithm.tile_size is not None:
            if len(algorithm.tile_size) != compiled.domain.ndim:
                raise ValueError("configured tile dimensionality does not match the lattice")
            tile_size = tuple(algorithm.tile_size) + (0,) * (3 - len(algorithm.tile_size))
        switching = algorithm.switching_interval or 0
    mechanics = []
    for component in components.mechanics:
        mechanics.append(_mechanics_entry(component, compiled, algorithm, columns))
    if len({m[2] for m in mechanics}) != len(mechanics):
        raise ValueError("mechanical components need distinct instance IDs")
    if len({m[0] for m in mechanics}) != len(mechanics):
        raise ValueError("mechanical components need distinct state properties")
    if J.shape[0] > MAX_TYPES:
        raise NotImplementedError("at most 16 cell types")
    if len(compiled.medium_ids) > MAX_MEDIA or len(table_owner) > MAX_MEDIA:
        raise NotImplementedError("at most 16 media")

    kinds, media = [], []
    for axis in range(3):
        if axis < domain.ndim:
            b = domain.boundaries[axis]
            (k0, m0), (k1, m1) = _face(b.negative), _face(b.positive)
        else:
            (k0, m0), (k1, m1) = (0, 0), (0, 0)
        kinds.append([k0, k1])
        media.append([m0, m1])
    dims = tuple(domain.dims) + (1,) * (3 - domain.ndim)
    has_obstacles = domain.has_obstacles
    return ModelSpec(
        algorithm=alg, ndim=domain.ndim, dims=dims, face_kind=kinds, face_medium=media,
        mutable_mask=domain.mutable_mask if has_obstacles else None,
        immutable_tags=domain.immutable_tags if has_obstacles else None,
        immutable_ids=domain.immutable_ids if has_obstacles else None,
        proposal=proposal_relation, contact=contact_rel, surface=compiled.tracker.relation,
        energy_order=order, volume_f64=bool(volume_f64), contact_f64=bool(contact_f64),
        boundary_f64=bool(boundary_f64), J=J, table_owner=tuple(table_owner),
        table_type=tuple(table_type), medium_ids=tuple(compiled.medium_ids),
        temperature=np.float32(algorithm.temperature), capacity=compiled.capacity,
        seed=int(seed), columns=columns, mechanics=tuple(mechanics), tile_size=tile_size,
        switching_interval=switching)


def state_arrays(compiled: CompiledScientificState, spec: ModelSpec):
    """Contiguous arrays in the order of cpm_upload_state / orc_state."""
    cap = compiled.capacity

    def col(name, dtype):
        key = spec.columns.get(name)
        if key is None:
            return np.zeros(cap, dtype=dtype)
        return np.ascontiguousarray(compiled.properties[key], dtype=dtype)

    return dict(
        tags=np.ascontiguousarray(compiled.tags, dtype=np.uint8),
        ids=np.ascontiguousarray(compiled.ids, dtype=np.uint32),
        finite_volumes=np.ascontiguousarray(compiled.finite_volumes, dtype=np.int32),
        medium_volumes=np.ascontiguousarray(
            np.resize(compiled.medium_volumes, max(len(compiled.medium_ids), 1)), dtype=np.int32),
        boundary_measures=np.ascontiguousarray(compiled.boundary_measures, dtype=np.int64),
        active=np.ascontiguousarray(compiled.active, dtype=np.uint8),
        cell_types=np.ascontiguousarray(compiled.cell_types, dtype=np.uint32),
        target_volume=col("target_volume", np.float64),
        volume_strength=col("volume_strength", np.float64),
        target_boundary=col("target_boundary", np.int64),
        boundary_strength=col("boundary_strength", np.float64),
        **{MECH_STATE_COLUMN[m[0]]: np.ascontiguousarray(compiled.properties[MECH_STATE_COLUMN[m[0]]],
                                                         dtype=np.float32)
           for m in spec.mechanics},
    )
