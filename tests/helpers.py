"""Shared builders for the parity tests.

phase13_fixture() rebuilds the tiny transition fixtures of the reference's Phase 13
evidence exactly as integration/transition/Phase13Fixtures.jl:121-190 does (Float32
spacing, relation weights, volume columns and contact matrix; energies (volume, contact);
BoundaryEdgeCount tracker only). scientific_fixture() restates the 4x4 / 4x4x4 fixture of
lib/CorePotts/test/test_scientific_hamiltonians.jl:1-45.
"""
import json
import os
from fractions import Fraction

import numpy as np

from corepotts import (AxisBoundary, BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain,
                       CellOwner, ClosedBoundary, ContactRole, LogicalPottsState, MediumOwner,
                       MediumTypeTable, PeriodicBoundary, ProposalRole, QuadraticBoundaryHamiltonian,
                       QuadraticVolumeHamiltonian, ScientificComponentSet, SurfaceRole,
                       UnorderedContactHamiltonian, build_spec, compile_scientific_state,
                       first_shell_relation, required_properties, state_arrays, static_relation,
                       von_neumann_offsets, moore_offsets)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_phase13():
    with open(os.path.join(GOLDEN, "phase13_transition_counts.json")) as handle:
        return json.load(handle)


def _rational(text) -> float:
    return float(Fraction(str(text)))


def parse_owner_encoding(encoding: str):
    owners = []
    for token in encoding.split(","):
        kind, ident = token.strip().split(":")
        owners.append(CellOwner(int(ident)) if kind == "cell" else MediumOwner(int(ident)))
    return owners


def phase13_fixture(row: dict, model: dict, algorithm):
    """Returns (compiled, proposal, components, algorithm_value, labels)."""
    dims = tuple(row["dimensions"])
    ndim = len(dims)
    f32 = np.float32
    spacing = tuple(f32(1.0) for _ in dims)
    offsets = von_neumann_offsets(ndim) if row["topology"] == "von_neumann" else moore_offsets(ndim)
    proposal = static_relation(ProposalRole, offsets, spacing=spacing)
    contact_rel = static_relation(ContactRole, offsets, spacing=spacing)
    surface = static_relation(SurfaceRole, offsets, spacing=spacing)
    bmap = {"periodic": PeriodicBoundary, "no_flux": ClosedBoundary}
    domain = CartesianDomain(dims, spacing=spacing,
                             boundaries=tuple(AxisBoundary(bmap[b]()) for b in row["boundaries"]))
    owners = parse_owner_encoding(row["source_state_encoding"])
    labels = []
    for o in owners:
        if o not in labels:
            labels.append(o)
    cell_ids = sorted({o for o in labels if o > 0})
    medium_ids = sorted({-o for o in labels if o < 0} | {model["medium_owner_id"]})
    for m in medium_ids:
        if -m not in labels:
            labels.append(-m)
    scale = _rational(row["energy_scale"])
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    base = np.array([_rational(v) for v in model["contact_matrix_base_row_major"]], dtype=np.float32)
    J = base.reshape(2, 2) * np.float32(scale)
    contact = UnorderedContactHamiltonian(J.astype(np.float32),
                                          MediumTypeTable({m: model["medium_type_id"] for m in medium_ids}),
                                          contact_rel)
    state = LogicalPottsState(np.array(owners, dtype=np.int64), model["cell_capacity"],
                              cell_types={c: model["cell_type_id"] for c in cell_ids},
                              medium_domains=medium_ids, property_schema=required_properties(volume),
                              dims=dims)
    for c in cell_ids:
        state.property_values("target_volume")[c - 1] = np.float32(_rational(model["target_volume"]))
        state.property_values("volume_strength")[c - 1] = np.float32(scale)
    tracker = BoundaryMeasureTracker(BoundaryEdgeCount(), surface)
    compiled = compile_scientific_state(state, domain, tracker)
    components = ScientificComponentSet(energies=(volume, contact))
    alg = algorithm(np.float32(_rational(row["temperature"])))
    return compiled, proposal, components, alg, labels


def scientific_fixture(dims, dtype=np.float32):
    """test_scientific_hamiltonians.jl:1-45 (volume, contact, boundary energies)."""
    ndim = len(dims)
    T = dtype
    spacing = tuple(T(1.0) for _ in dims)
    volume = QuadraticVolumeHamiltonian(number_type=T)
    surface = first_shell_relation(SurfaceRole, ndim, spacing=spacing)
    boundary = QuadraticBoundaryHamiltonian(BoundaryEdgeCount(), surface, number_type=T)
    schema = {}
    schema.update(required_properties(volume))
    schema.update(required_properties(boundary))
    owners = np.full(dims, MediumOwner(1), dtype=np.int64)
    if ndim == 2:
        owners[0, 0] = owners[1, 0] = owners[0, 1] = CellOwner(1)
        owners[-1, -1] = owners[-2, -1] = owners[-1, -2] = CellOwner(2)
    else:
        owners[0, 0, 0] = owners[1, 0, 0] = owners[0, 1, 0] = CellOwner(1)
        owners[-1, -1, -1] = owners[-2, -1, -1] = owners[-1, -2, -1] = CellOwner(2)
    state = LogicalPottsState(owners, 2, cell_types={1: 2, 2: 2}, medium_domains=(1,),
                              property_schema=schema)
    state.property_values("target_volume")[:] = T(4)
    state.property_values("volume_strength")[:] = T(1.25)
    state.property_values("target_boundary")[:] = 8
    state.property_values("boundary_strength")[:] = T(0.75)
    domain = CartesianDomain(dims, spacing=spacing)
    proposal = first_shell_relation(ProposalRole, ndim, spacing=domain.spacing)
    contact_rel = first_shell_relation(ContactRole, ndim, spacing=domain.spacing)
    contact = UnorderedContactHamiltonian(np.array([[0, 3], [3, 1]], dtype=T),
                                          MediumTypeTable({1: 1}), contact_rel)
    tracker = BoundaryMeasureTracker(BoundaryEdgeCount(), surface)
    compiled = compile_scientific_state(state, domain, tracker)
    components = ScientificComponentSet(energies=(volume, contact, boundary))
    return compiled, proposal, components


def oracle_for(compiled, proposal, components, algorithm, seed, alg_override=None, initialize_mechanics=True):
    """Oracle integrator as init_scientific builds it (mechanical states initialized)."""
    from oracle.bridge import OracleSim
    spec = build_spec(compiled, proposal, components, algorithm, seed)
    sim = OracleSim(spec, state_arrays(compiled, spec), alg_override)
    if spec.mechanics and initialize_mechanics:
        sim.initialize_mechanics()
    return sim, spec


# ---------------------------------------------------------------------------
# The reference's realistic single-cell-migration evidence
# ---------------------------------------------------------------------------
def load_migration():
    with open(os.path.join(GOLDEN, "realistic_single_cell_migration.json")) as handle:
        return json.load(handle)


def cell_center(owner_words, dims):
    """_cell_center (integration/transition/RealisticScaleRunner.jl:230-241): per axis the
    circular mean of the 1-based coordinates of the cell's sites."""
    L0, L1 = dims
    idx = np.nonzero(owner_words == 1)[0]
    coords = (idx % L0 + 1, idx // L0 + 1)
    out = []
    for axis, d in enumerate(dims):
        angles = 2 * np.pi * (coords[axis] - 1) / d
        angle = np.arctan2(np.mean(np.sin(angles)), np.mean(np.cos(angles)))
        out.append(1 + np.mod(angle, 2 * np.pi) * d / (2 * np.pi))
    return tuple(out)


def _wrapped(first, last, dims):
    out = []
    for axis, d in enumerate(dims):
        delta = last[axis] - first[axis]
        if delta > d / 2:
            delta -= d
        if delta < -d / 2:
            delta += d
        out.append(delta)
    return out


def _mixing_time(energies, cadence, window):
    """_mixing_time (RealisticScaleRunner.jl:252-264)."""
    if not energies:
        return 0
    window = min(window, len(energies))
    rolling = [float(np.mean(energies[max(0, i - window + 1):i + 1])) for i in range(len(energies))]
    target = rolling[-1]
    tol = 0.05 * max(abs(target), np.finfo(np.float64).eps)
    for i in range(len(rolling)):
        if all(abs(v - target) <= tol for v in rolling[i:]):
            return (i + 1) * cadence
    return len(energies) * cadence


def migration_summary(energies, boundaries, centers, dims=(64, 64), cadence=5, window=20):
    """The observables run_realistic_replica (RealisticScaleRunner.jl:266-360) records for the
    single-cell workload, from the per-observation energy per site, boundary measure and
    cell centre."""
    disp = _wrapped(centers[0], centers[-1], dims)
    steps = [_wrapped(centers[i - 1], centers[i], dims) for i in range(1, len(centers))]
    distances = [float(np.hypot(*s)) for s in steps]
    net = float(np.hypot(*disp))
    return {"energy_per_mutable_site_terminal": energies[-1],
            "energy_per_mutable_site_mean": float(np.mean(energies)),
            "boundary_measure_mean": float(np.mean(boundaries)),
            "mixing_time_mcs": _mixing_time(energies, cadence, window),
            "net_displacement": net,
            "mean_squared_displacement": float(np.mean([s[0] ** 2 + s[1] ** 2 for s in steps])),
            "speed": float(np.mean(distances)) / cadence,
            "persistence": 0.0 if sum(distances) == 0 else net / sum(distances)}


def migration_observables(step, observe, dims=(64, 64), burn_in=500, measured=2000, cadence=5,
                          window=20):
    """run_realistic_replica (RealisticScaleRunner.jl:266-360) for the single-cell workload:
    `step(n)` advances n MCS; `observe()` returns (global energy, boundary measure of cell 1,
    owner words) after a step."""
    step(burn_in)
    energies, boundaries, centers = [], [], []
    sites = dims[0] * dims[1]
    for _ in range(measured // cadence):
        step(cadence)
        energy, boundary, owners = observe()
        energies.append(energy / sites)
        boundaries.append(float(boundary))
        centers.append(cell_center(owners, dims))
    return migration_summary(energies, boundaries, centers, dims, cadence, window)


EXACT_MIGRATION_FIELDS = ("energy_per_mutable_site_terminal", "energy_per_mutable_site_mean",
                          "boundary_measure_mean", "mixing_time_mcs")
FLOAT_MIGRATION_FIELDS = ("net_displacement", "mean_squared_displacement", "speed", "persistence")


def oracle_migration_replica(algorithm: str, dynamics_seed: int):
    """One replica of the single-cell-migration workload on the oracle (algorithm
    'sequential' or 'checkerboard')."""
    from corepotts import workloads
    from oracle.bridge import OracleSim
    model = workloads.single_cell_migration("checkerboard", dynamics_seed)
    spec = build_spec(*model)
    sim = OracleSim(spec, state_arrays(model[0], spec), 0 if algorithm == "sequential" else None)

    def observe():
        st = sim.store()
        words = np.where(st["tags"] == 1, st["ids"].astype(np.int64), -st["ids"].astype(np.int64))
        return sim.total_energy(), int(st["boundary_measures"][0]), words

    return migration_observables(sim.step, observe)


# ---------------------------------------------------------------------------
# Deep-run digests (tests/golden/make_deep_digests.py writes them from the oracle)
# ---------------------------------------------------------------------------
DIGEST_FIELDS = (("tags", np.uint8), ("ids", np.uint32), ("finite_volumes", np.int32),
                 ("medium_volumes", np.int32), ("boundary_measures", np.int64))


def state_digest(state: dict) -> str:
    """BLAKE2b-128 of the ownership and tracker arrays of a downloaded or oracle state."""
    import hashlib
    h = hashlib.blake2b(digest_size=16)
    for key, dtype in DIGEST_FIELDS:
        h.update(np.ascontiguousarray(state[key], dtype=dtype).tobytes())
    return h.hexdigest()


def load_deep_digests():
    with open(os.path.join(GOLDEN, "deep_digests.json")) as handle:
        return json.load(handle)


def deep_serving_jobs(n2d: int = 6, n3d: int = 2, mcs: int = 20):
    """Jobs of the served stream at their configured sizes: the first n2d 2D jobs of side at
    least 1024 and n3d 3D jobs of side at least 96, each run for `mcs` MCS."""
    from corepotts import workloads
    picked2, picked3 = [], []
    for job in workloads.serving_jobs(400):
        if job.ndim == 2 and job.side >= 1024 and len(picked2) < n2d:
            picked2.append(job._replace(mcs=mcs))
        if job.ndim == 3 and job.side >= 96 and len(picked3) < n3d:
            picked3.append(job._replace(mcs=mcs))
    return picked2 + picked3
