"""Host-side mirror of the CorePotts model types used by the Metropolis hot path.

Names, argument meaning and error behaviour follow the reference package (paths relative
to the reference checkout):

* owners, ``CellOwner``/``MediumOwner``        lib/CorePotts/src/state/logical.jl:9-45
* ``CartesianDomain`` and boundary kinds       lib/CorePotts/src/spatial/cartesian.jl:1-122
* ``static_relation`` canonical order           lib/CorePotts/src/spatial/cartesian.jl:290-350
* ``first_shell_relation``                      lib/CorePotts/src/spatial/cartesian.jl:352-359
* energy components and ``MediumTypeTable``     lib/CorePotts/src/components/scientific_components.jl
* ``BoundaryMeasureTracker``, tracker rebuild   lib/CorePotts/src/components/scientific_trackers.jl
* ``compile_scientific_state``                  lib/CorePotts/src/components/scientific_trackers.jl:210-228
* algorithm types                               lib/CorePotts/src/algorithms/sequential.jl:1-56

Owners are carried as integer owner words: ``+id`` for a cell and ``-id`` for a medium.
Lattices are flat arrays in the reference's column-major site order (x fastest), so the
0-based index of site ``s`` is ``s - 1``.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from typing import Dict, Iterable, Optional, Sequence, Tuple

import numpy as np

CELL_TAG = 1
MEDIUM_TAG = 2


def CellOwner(cell_id: int) -> int:
    """Owner word of a finite cell (logical.jl:29)."""
    cell_id = int(cell_id)
    if cell_id <= 0:
        raise ValueError("owner identity must be positive")
    return cell_id


def MediumOwner(medium_id: int) -> int:
    """Owner word of a conceptual medium domain (logical.jl:33)."""
    medium_id = int(medium_id)
    if medium_id <= 0:
        raise ValueError("owner identity must be positive")
    return -medium_id


def is_cell_owner(word: int) -> bool:
    return word > 0


def owner_tags_ids(words: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Split owner words into the reference's SoA (tags::UInt8, ids::UInt32)."""
    words = np.asarray(words, dtype=np.int64)
    tags = np.where(words > 0, CELL_TAG, MEDIUM_TAG).astype(np.uint8)
    ids = np.abs(words).astype(np.uint32)
    return tags, ids


def owner_words(tags: np.ndarray, ids: np.ndarray) -> np.ndarray:
    tags = np.asarray(tags)
    ids = np.asarray(ids).astype(np.int64)
    return np.where(tags == CELL_TAG, ids, -ids)


# ---------------------------------------------------------------------------
# Domain
# ---------------------------------------------------------------------------
class PeriodicBoundary:
    kind = 0


class ClosedBoundary:
    kind = 1


class FixedExterior:
    kind = 2

    def __init__(self, owner: int):
        if owner >= 0:
            raise ValueError("fixed exterior owners must be conceptual medium or wall domains")
        self.owner = int(owner)


class AxisBoundary:
    def __init__(self, negative, positive=None):
        self.negative = negative
        self.positive = negative if positive is None else positive


def _dtype_of(values) -> np.dtype:
    return np.result_type(*[np.asarray(v).dtype if isinstance(v, np.generic) else np.float64
                            for v in values])


class CartesianDomain:
    """Rectangular ownership domain (cartesian.jl:40-122)."""

    def __init__(self, dims: Sequence[int], spacing=None, boundaries=None, obstacles=()):
        dims = tuple(int(d) for d in dims)
        if len(dims) not in (2, 3):
            raise ValueError("Cartesian ownership domains support two or three dimensions")
        if any(d <= 0 for d in dims):
            raise ValueError("Cartesian dimensions must be positive")
        n = int(np.prod(dims, dtype=np.int64))
        if n > 0xFFFFFFFF:
            raise ValueError("Cartesian site count must fit the compiled UInt32 site representation")
        if spacing is None:
            spacing = tuple(np.float64(1.0) for _ in dims)
        if len(spacing) != len(dims):
            raise ValueError("spacing dimension mismatch")
        self.spacing_dtype = _dtype_of(spacing)
        self.spacing = tuple(self.spacing_dtype.type(s) for s in spacing)
        if boundaries is None:
            boundaries = tuple(AxisBoundary(PeriodicBoundary()) for _ in dims)
        if len(boundaries) != len(dims):
            raise ValueError("one boundary pair is required per axis")
        for b in boundaries:
            if isinstance(b.negative, PeriodicBoundary) != isinstance(b.positive, PeriodicBoundary):
                raise ValueError("periodic boundary faces must occur as a paired axis")
        self.dims = dims
        self.boundaries = tuple(boundaries)
        self.mutable_mask = np.ones(n, dtype=np.uint8)
        self.immutable_tags = np.zeros(n, dtype=np.uint8)
        self.immutable_ids = np.zeros(n, dtype=np.uint32)
        for site, owner in obstacles:
            index = self._site_index(site)
            if owner >= 0:
                raise ValueError("obstacle owners must be conceptual medium or wall domains")
            if not self.mutable_mask[index]:
                raise ValueError("an obstacle site may be declared only once")
            self.mutable_mask[index] = 0
            self.immutable_tags[index] = MEDIUM_TAG
            self.immutable_ids[index] = -owner
        if not self.mutable_mask.any():
            raise ValueError("a Cartesian ownership domain requires a mutable site")

    def _site_index(self, site) -> int:
        if isinstance(site, tuple):
            idx = 0
            stride = 1
            for c, d in zip(site, self.dims):
                if not 1 <= c <= d:
                    raise IndexError("obstacle outside the domain")
                idx += (c - 1) * stride
                stride *= d
            return idx
        if not 1 <= site <= self.n_sites:
            raise IndexError("obstacle outside the domain")
        return int(site) - 1

    @property
    def ndim(self) -> int:
        return len(self.dims)

    @property
    def n_sites(self) -> int:
        return int(np.prod(self.dims, dtype=np.int64))

    @property
    def has_obstacles(self) -> bool:
        return not bool(self.mutable_mask.all())

    def mutable_site_count(self) -> int:
        return int(self.mutable_mask.sum())


# ---------------------------------------------------------------------------
# Relations
# ---------------------------------------------------------------------------
class _Role:
    name = ""
    allows_fixed = False


class ProposalRole(_Role):
    name = "ProposalRole"


class ContactRole(_Role):
    name = "ContactRole"
    allows_fixed = True


class SurfaceRole(_Role):
    name = "SurfaceRole"
    allows_fixed = True


@dataclass(frozen=True)
class StaticCartesianRelation:
    role: type
    offsets: Tuple[Tuple[int, ...], ...]
    weights: Tuple[float, ...]
    weight_dtype: np.dtype
    opposite: Tuple[int, ...]
    symmetric: bool = True

    def direction_count(self) -> int:
        return len(self.offsets)


def static_relation(role, offsets: Iterable[Sequence[int]], spacing=None, weights=None,
                    symmetric: bool = True) -> StaticCartesianRelation:
    """Canonical relation: directions sorted by (physical squared length, offset tuple)
    (cartesian.jl:290-350)."""
    raw = [tuple(int(c) for c in o) for o in offsets]
    if not raw:
        raise ValueError("a Cartesian relation requires at least one offset")
    n = len(raw[0])
    if n not in (2, 3) or any(len(o) != n for o in raw):
        raise ValueError("Cartesian relation offsets must be two- or three-dimensional")
    if any(all(c == 0 for c in o) for o in raw):
        raise ValueError("Cartesian relation offsets must be nonzero")
    if len(set(raw)) != len(raw):
        raise ValueError("duplicate Cartesian relation offsets are invalid")
    if spacing is None:
        spacing = tuple(np.float64(1.0) for _ in range(n))
    sdtype = _dtype_of(spacing)
    phys = np.array([sdtype.type(s) for s in spacing], dtype=sdtype)
    if weights is None:
        wvals = [sdtype.type(1.0)] * len(raw)
        wdtype = sdtype
    else:
        weights = list(weights)
        if len(weights) != len(raw):
            raise ValueError("one relation weight is required per offset")
        wdtype = np.result_type(sdtype, _dtype_of(weights))
        wvals = [wdtype.type(w) for w in weights]
    wdtype = np.result_type(sdtype, wdtype)
    if any(not np.isfinite(w) or w < 0 for w in wvals):
        raise ValueError("relation weights must be finite and non-negative")

    def key(i):
        o = np.array(raw[i], dtype=sdtype)
        return (float(np.sum((o * phys) ** 2, dtype=sdtype)), raw[i])

    order = sorted(range(len(raw)), key=key)
    canonical = tuple(raw[i] for i in order)
    cweights = tuple(float(wdtype.type(wvals[i])) for i in order)
    opposite = []
    for o in canonical:
        neg = tuple(-c for c in o)
        opposite.append(canonical.index(neg) + 1 if neg in canonical else 0)
    if symmetric:
        if any(x == 0 for x in opposite):
            raise ValueError("a symmetric relation requires every opposite offset")
        if any(cweights[i] != cweights[opposite[i] - 1] for i in range(len(canonical))):
            raise ValueError("a symmetric relation requires equal weights for opposite offsets")
    return StaticCartesianRelation(role, canonical, cweights, np.dtype(wdtype), tuple(opposite),
                                   symmetric)


def von_neumann_offsets(ndim: int):
    neg = [tuple(-1 if a == d else 0 for a in range(ndim)) for d in range(ndim)]
    pos = [tuple(1 if a == d else 0 for a in range(ndim)) for d in range(ndim)]
    return neg + pos


def moore_offsets(ndim: int):
    # Iterators.product(-1:1, ...) enumerates the first axis fastest
    out = []
    for rev in itertools.product((-1, 0, 1), repeat=ndim):
        o = tuple(reversed(rev))
        if any(o):
            out.append(o)
    return out


def first_shell_relation(role, ndim: int, spacing=None, weights=None) -> StaticCartesianRelation:
    return static_relation(role, von_neumann_offsets(ndim), spacing=spacing, weights=weights)


def moore_relation(role, ndim: int, spacing=None) -> StaticCartesianRelation:
    return static_relation(role, moore_offsets(ndim), spacing=spacing)


def validate_relation_domain(domain: CartesianDomain, relation: StaticCartesianRelation):
    """Reject self-edges and periodic aliases (cartesian.jl:567-603)."""
    if len(relation.offsets[0]) != domain.ndim:
        raise ValueError("relation and Cartesian domain dimensions must match")
    reach = max(abs(c) for o in relation.offsets for c in o)
    wraps = [isinstance(b.negative, PeriodicBoundary) for b in domain.boundaries]
    # Distinct offsets within `reach` can only alias, or an offset realize its own site,
    # through a periodic axis no longer than 2 * reach.
    if all(not w or d > 2 * reach for w, d in zip(wraps, domain.dims)):
        return relation
    n = domain.n_sites
    idx = np.arange(n, dtype=np.int64)
    realized = []
    for off in relation.offsets:
        ok, lin = _realize_index(domain, off)
        if np.any(ok & (lin == idx)):
            raise ValueError("a nonzero relation offset realizes a self-edge")
        realized.append((ok, lin))
    for a in range(len(realized)):
        for b in range(a + 1, len(realized)):
            oka, la = realized[a]
            okb, lb = realized[b]
            if np.any(oka & okb & (la == lb)):
                raise ValueError("distinct relation offsets realize the same neighbor")
    return relation


def _realize_index(domain: CartesianDomain, offset: Sequence[int]):
    """Realized in-domain linear index per site (False where closed, fixed or invalid)."""
    n = domain.n_sites
    rem = np.arange(n, dtype=np.int64)
    ok = np.ones(n, dtype=bool)
    lin = np.zeros(n, dtype=np.int64)
    stride = 1
    for axis, d in enumerate(domain.dims):
        c = rem % d
        rem = rem // d
        v = c + int(offset[axis])
        out = (v < 0) | (v >= d)
        if isinstance(domain.boundaries[axis].negative, PeriodicBoundary):
            v = np.mod(v, d)
        else:
            ok &= ~out
            v = np.clip(v, 0, d - 1)
        lin += v * stride
        stride *= d
    return ok, lin


# ---------------------------------------------------------------------------
# Components
# ---------------------------------------------------------------------------
class BoundaryEdgeCount:
    """Exact number of outward unlike-owner incidences (scientific_components.jl:343-344)."""


@dataclass
class BoundaryMeasureTracker:
    metric: object
    relation: StaticCartesianRelation


class MediumTypeTable:
    """Static mapping medium owner id -> cell-type id, sorted by owner id."""

    def __init__(self, entries: Dict[int, int]):
        items = sorted((int(k), int(v)) for k, v in dict(entries).items())
        self.owner_ids = tuple(k for k, _ in items)
        self.type_ids = tuple(v for _, v in items)


@dataclass
class QuadraticVolumeHamiltonian:
    number_type: type = np.float64
    target: str = "target_volume"
    strength: str = "volume_strength"


@dataclass
class UnorderedContactHamiltonian:
    interactions: np.ndarray
    medium_types: MediumTypeTable
    relation: StaticCartesianRelation

    def __post_init__(self):
        J = np.asarray(self.interactions)
        if J.ndim != 2 or J.shape[0] != J.shape[1] or J.shape[0] == 0:
            raise ValueError("contact interactions must form a nonempty square matrix")
        if not np.all(np.isfinite(J)):
            raise ValueError("contact interactions must be finite")
        if not np.array_equal(J, J.T):
            raise ValueError("unordered contact interactions must be symmetric")
        if any(not 1 <= t <= J.shape[0] for t in self.medium_types.type_ids):
            raise ValueError("a medium type lies outside the interaction matrix")
        if J.dtype not in (np.float32, np.float64):
            J = J.astype(np.float64)
        self.interactions = J


@dataclass
class QuadraticBoundaryHamiltonian:
    metric: object
    relation: StaticCartesianRelation
    number_type: type = np.float64
    target: str = "target_boundary"
    strength: str = "boundary_strength"


# ---------------------------------------------------------------------------
# Mechanics (scientific_components.jl:130-240, 575-660)
# ---------------------------------------------------------------------------
MECH_PRESSURE, MECH_TENSION = 1, 2
MECH_STATE_COLUMN = {MECH_PRESSURE: "volume_pressure", MECH_TENSION: "surface_tension"}
# MechanicalInitialization (scientific_components.jl:130-134)
ConstitutiveMeanInitialization = 1
StationaryMechanicalInitialization = 2
PreserveMechanicalInitialization = 3


class AlgorithmTemperatureNoise:
    """Mechanical noise scale = the algorithm temperature (scientific_components.jl:136-137)."""


class FixedMechanicalNoise:
    """Mechanical noise scale independent of the acceptance temperature (:139-150)."""

    def __init__(self, value: float):
        if not np.isfinite(value) or value < 0:
            raise ValueError("fixed mechanical noise must be finite and non-negative")
        self.value = float(value)


class _Mechanics:
    def __init__(self, eta, noise, initialization, instance_id, number_type, target, strength, state,
                 what):
        rate = np.dtype(number_type).type(eta)
        if not (np.isfinite(rate) and rate > 0):
            raise ValueError(f"{what} relaxation rate must be finite and positive")
        if isinstance(noise, AlgorithmTemperatureNoise):
            noise = None
        if noise is not None and not isinstance(noise, FixedMechanicalNoise):
            raise ValueError("unsupported mechanical-noise declaration")
        if initialization not in (ConstitutiveMeanInitialization, StationaryMechanicalInitialization,
                                  PreserveMechanicalInitialization):
            raise ValueError("unknown mechanical initialization")
        if not 0 < int(instance_id) <= 0xFFFF:
            raise ValueError("mechanical component instance ID must be positive and fit UInt16")
        if len({target, strength, state}) != 3:
            raise ValueError(f"{what} target, strength, and state properties must be distinct")
        self.eta = rate
        self.noise = noise                   # None: AlgorithmTemperatureNoise
        self.initialization = initialization
        self.instance_id = int(instance_id)
        self.number_type = np.dtype(number_type).type
        self.target, self.strength, self.state = target, strength, state


class FluctuatingVolumePressure(_Mechanics):
    """Per-cell pressure with an exact frozen-volume OU transition (scientific_components.jl:152-182)."""

    def __init__(self, eta=1.0, noise=None, initialization=ConstitutiveMeanInitialization, instance_id=1,
                 number_type=np.float32, target="target_volume", strength="volume_strength",
                 state="volume_pressure"):
        super().__init__(eta, noise, initialization, instance_id, number_type, target, strength, state,
                         "volume-pressure")


class FluctuatingSurfaceTension(_Mechanics):
    """Per-cell tension with an exact frozen-surface OU transition (scientific_components.jl:576-616)."""

    def __init__(self, metric, relation, eta=1.0, noise=None, initialization=ConstitutiveMeanInitialization,
                 instance_id=1, number_type=np.float32, target="target_boundary",
                 strength="boundary_strength", state="surface_tension"):
        super().__init__(eta, noise, initialization, instance_id, number_type, target, strength, state,
                         "surface-tension")
        self.metric = metric
        self.relation = relation


@dataclass
class ScientificComponentSet:
    energies: Tuple = ()
    drives: Tuple = ()
    constraints: Tuple = ()
    kinetic_modifiers: Tuple = ()
    mechanics: Tuple = ()


def required_properties(component) -> Dict[str, np.dtype]:
    if isinstance(component, QuadraticVolumeHamiltonian):
        t = np.dtype(component.number_type)
        return {component.target: t, component.strength: t}
    if isinstance(component, QuadraticBoundaryHamiltonian):
        t = np.dtype(component.number_type)
        tt = np.dtype(np.int64) if isinstance(component.metric, BoundaryEdgeCount) else t
        return {component.target: tt, component.strength: t}
    if isinstance(component, FluctuatingVolumePressure):
        t = np.dtype(component.number_type)
        return {component.target: t, component.strength: t, component.state: t}
    if isinstance(component, FluctuatingSurfaceTension):
        t = np.dtype(component.number_type)
        tt = np.dtype(np.int64) if isinstance(component.metric, BoundaryEdgeCount) else t
        return {component.target: tt, component.strength: t, component.state: t}
    return {}


# ---------------------------------------------------------------------------
# Algorithms (sequential.jl:1-56)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class _Algorithm:
    temperature: np.float32 = np.float32(20.0)

    def __post_init__(self):
        t = self.temperature
        if not isinstance(t, np.floating):
            t = np.float32(t)
        if not np.isfinite(t) or t < 0:
            raise ValueError("temperature must be finite and non-negative")
        object.__setattr__(self, "temperature", t)


class CheckerboardSweepCPM(_Algorithm):
    pass


class LotteryCPM(_Algorithm):
    pass


class SequentialCPM(_Algorithm):
    pass


class TiledCheckerboardCPM:
    """Deterministic tile-local CPM with snapshot and reconciliation subrounds
    (tiled_checkerboard.jl:15-47): tile_size None (4 x 4 / 2 x 2 x 2) or a 2- or 3-tuple of
    positive extents, switching_interval None or a positive UInt16, shared_memory one of
    "auto", "required", "disabled" (launch geometry only)."""

    def __init__(self, temperature=np.float32(20.0), tile_size=None, switching_interval=None,
                 shared_memory="auto"):
        t = temperature if isinstance(temperature, np.floating) else np.float32(temperature)
        if not np.isfinite(t) or t < 0:
            raise ValueError("tiled checkerboard temperature must be finite and non-negative")
        if tile_size is not None:
            if not isinstance(tile_size, tuple) or len(tile_size) not in (2, 3):
                raise ValueError("tile_size must be nothing or a two-/three-dimensional tuple")
            if not all(isinstance(v, (int, np.integer)) and v > 0 for v in tile_size):
                raise ValueError("tile dimensions must be positive integers")
            tile_size = tuple(int(v) for v in tile_size)
        if switching_interval is not None:
            if not isinstance(switching_interval, (int, np.integer)) or switching_interval <= 0:
                raise ValueError("switching_interval must be nothing or a positive integer")
            if switching_interval > 0xFFFF:
                raise ValueError("switching_interval must fit UInt16")
            switching_interval = int(switching_interval)
        if shared_memory not in ("auto", "required", "disabled"):
            raise ValueError("shared_memory must be :auto, :required, or :disabled")
        self.temperature = t
        self.tile_size = tile_size
        self.switching_interval = switching_interval
        self.shared_memory = shared_memory


# ---------------------------------------------------------------------------
# State
# ---------------------------------------------------------------------------
class LogicalPottsState:
    """Host authority for one lattice (logical.jl:107-124, 327)."""

    def __init__(self, owners, capacity: int, cell_types: Optional[Dict[int, int]] = None,
                 medium_domains: Sequence[int] = (), property_schema: Optional[Dict] = None,
                 dims: Optional[Sequence[int]] = None, active_ids=None):
        arr = np.asarray(owners)
        if dims is None:
            dims = arr.shape
            flat = arr.reshape(-1, order="F")
        else:
            flat = arr.reshape(-1)
        self.dims = tuple(int(d) for d in dims)
        self.owners = flat.astype(np.int64)
        self.capacity = int(capacity)
        if self.capacity < 1:
            raise ValueError("capacity must be positive")
        cells = self.owners[self.owners > 0]
        if cells.size and cells.max() > self.capacity:
            raise ValueError("a cell owner exceeds the capacity")
        self.medium_ids = tuple(int(m) for m in medium_domains)
        media = -self.owners[self.owners < 0]
        for m in np.unique(media):
            if int(m) not in self.medium_ids:
                raise ValueError("every medium owner must be a declared medium domain")
        self.cell_types = np.zeros(self.capacity, dtype=np.uint32)
        for cid, tid in (cell_types or {}).items():
            self.cell_types[int(cid) - 1] = int(tid)
        self.active = np.zeros(self.capacity, dtype=np.uint8)
        if active_ids is None:
            present = np.unique(cells)
            self.active[present - 1] = 1
        else:
            for cid in active_ids:
                self.active[int(cid) - 1] = 1
        self.properties: Dict[str, np.ndarray] = {}
        for name, dtype in (property_schema or {}).items():
            self.properties[name] = np.zeros(self.capacity, dtype=dtype)

    def property_values(self, name: str) -> np.ndarray:
        return self.properties[name]

    def finite_volumes(self) -> np.ndarray:
        cells = self.owners[self.owners > 0]
        return np.bincount(cells - 1, minlength=self.capacity).astype(np.int32)

    def medium_volumes(self) -> np.ndarray:
        out = np.zeros(len(self.medium_ids), dtype=np.int32)
        for i, m in enumerate(self.medium_ids):
            out[i] = int(np.count_nonzero(self.owners == -m))
        return out


def _realize_all(domain: CartesianDomain, offset: Sequence[int], owners: np.ndarray):
    """Vectorized neighbour realization of every site for contact/surface roles.

    Returns (valid mask, neighbour owner words) following cartesian.jl:445-525."""
    dims = domain.dims
    n = domain.n_sites
    if all(isinstance(b.negative, PeriodicBoundary) and isinstance(b.positive, PeriodicBoundary)
           for b in domain.boundaries):
        # every axis wraps: the neighbour of x is x + offset mod dims, a cyclic shift
        grid = owners.reshape(dims, order="F")
        shifted = np.roll(grid, shift=[-int(offset[a]) for a in range(len(dims))],
                          axis=tuple(range(len(dims))))
        return np.ones(n, dtype=bool), shifted.reshape(-1, order="F")
    idx = np.arange(n, dtype=np.int64)
    coords = []
    rem = idx
    for d in dims:
        coords.append(rem % d)
        rem = rem // d
    closed = np.zeros(n, dtype=bool)
    fixed = np.zeros(n, dtype=bool)
    invalid = np.zeros(n, dtype=bool)
    fixed_owner = np.zeros(n, dtype=np.int64)
    new = []
    for axis, (c, d) in enumerate(zip(coords, dims)):
        v = c + int(offset[axis])
        b = domain.boundaries[axis]
        for face, sel in ((b.negative, v < 0), (b.positive, v >= d)):
            if not sel.any():
                continue
            if isinstance(face, PeriodicBoundary):
                v = np.where(sel, np.mod(v, d), v)
            elif isinstance(face, ClosedBoundary):
                invalid |= sel & fixed
                closed |= sel
            else:
                incompatible = sel & fixed & (fixed_owner != face.owner)
                invalid |= sel & (closed | incompatible)
                fixed |= sel
                fixed_owner = np.where(sel, face.owner, fixed_owner)
        new.append(v)
    ok = ~(closed | invalid)
    lin = np.zeros(n, dtype=np.int64)
    stride = 1
    for v, d in zip(new, dims):
        lin += np.clip(v, 0, d - 1) * stride
        stride *= d
    words = np.where(fixed, fixed_owner, owners[lin])
    return ok, words


def rebuild_boundary_measures(state: LogicalPottsState, domain: CartesianDomain,
                              relation: StaticCartesianRelation) -> np.ndarray:
    """BoundaryEdgeCount tracker rebuilt from ownership (scientific_trackers.jl:33-42)."""
    owners = state.owners
    result = np.zeros(state.capacity, dtype=np.int64)
    mutable = domain.mutable_mask.astype(bool)
    for off in relation.offsets:
        ok, words = _realize_all(domain, off, owners)
        sel = ok & mutable & (owners > 0) & (words != owners)
        result += np.bincount(owners[sel] - 1, minlength=state.capacity).astype(np.int64)
    result[state.active == 0] = 0
    return result


@dataclass
class CompiledScientificState:
    """Compiled ownership, domain, trackers and property columns (scientific_trackers.jl:70-78)."""
    domain: CartesianDomain
    tracker: BoundaryMeasureTracker
    capacity: int
    medium_ids: Tuple[int, ...]
    tags: np.ndarray
    ids: np.ndarray
    finite_volumes: np.ndarray
    medium_volumes: np.ndarray
    boundary_measures: np.ndarray
    active: np.ndarray
    cell_types: np.ndarray
    properties: Dict[str, np.ndarray] = field(default_factory=dict)

    @property
    def owners(self) -> np.ndarray:
        return owner_words(self.tags, self.ids)

    def copy(self) -> "CompiledScientificState":
        return CompiledScientificState(
            self.domain, self.tracker, self.capacity, self.medium_ids, self.tags.copy(),
            self.ids.copy(), self.finite_volumes.copy(), self.medium_volumes.copy(),
            self.boundary_measures.copy(), self.active.copy(), self.cell_types.copy(),
            {k: v.copy() for k, v in self.properties.items()})


def compile_scientific_state(state: LogicalPottsState, domain: CartesianDomain,
                             tracker: BoundaryMeasureTracker) -> CompiledScientificState:
    if tuple(state.dims) != tuple(domain.dims):
        raise ValueError("logical state and ownership domain dimensions must match")
    if not isinstance(tracker.metric, BoundaryEdgeCount):
        raise NotImplementedError("only the BoundaryEdgeCount tracker is compiled for this engine")
    immutable = domain.mutable_mask == 0
    if immutable.any():
        imm_words = owner_words(domain.immutable_tags[immutable], domain.immutable_ids[immutable])
        if not np.array_equal(imm_words, state.owners[immutable]):
            raise ValueError("logical ownership does not match the immutable obstacle owner")
    validate_relation_domain(domain, tracker.relation)
    tags, ids = owner_tags_ids(state.owners)
    return CompiledScientificState(
        domain=domain, tracker=tracker, capacity=state.capacity, medium_ids=state.medium_ids,
        tags=tags, ids=ids, finite_volumes=state.finite_volumes(),
        medium_volumes=state.medium_volumes(),
        boundary_measures=rebuild_boundary_measures(state, domain, tracker.relation),
        active=state.active.copy(), cell_types=state.cell_types.copy(),
        properties={k: v.copy() for k, v in state.properties.items()})


def tracker_conformance_errors(compiled: CompiledScientificState) -> list:
    """Trackers recomputed from ownership (scientific_trackers.jl:399-423)."""
    logical = LogicalPottsState(compiled.owners, compiled.capacity, medium_domains=compiled.medium_ids,
                                dims=compiled.domain.dims)
    logical.active = compiled.active.copy()
    errors = []
    if not np.array_equal(logical.finite_volumes(), compiled.finite_volumes):
        errors.append("finite-cell volume tracker differs from ownership recomputation")
    if not np.array_equal(logical.medium_volumes(), compiled.medium_volumes):
        errors.append("medium volume tracker differs from ownership recomputation")
    rebuilt = rebuild_boundary_measures(logical, compiled.domain, compiled.tracker.relation)
    expected = np.where(compiled.active == 1, rebuilt, compiled.boundary_measures)
    if not np.array_equal(expected, compiled.boundary_measures):
        errors.append("boundary tracker differs from ownership recomputation")
    return errors
