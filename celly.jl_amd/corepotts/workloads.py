"""Deterministic workload builders: square/cube cell layouts on periodic lattices.

Each builder returns (compiled_state, proposal_relation, components, algorithm, seed) ready
for ``init_scientific``. Layout rules:

* plain_2d: L x L periodic lattice, medium id 1 (type 1); cells are side x side squares whose
  lower-left corners sit on a regular grid; cell id k (1-based) takes type 2 + (k-1) mod ntypes.
* The contact matrix over types (1 = medium, 2, 3) is J11=0, J12=J13=16, J22=2, J33=14, J23=11.
* Volume constraint lambda_V, target V*; optional boundary constraint lambda_S, target S* on
  the von Neumann edge count; Moore contact relation; von Neumann proposal relation.
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np

from .model import (BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain,
                    CheckerboardSweepCPM, LogicalPottsState, LotteryCPM, MediumTypeTable,
                    ProposalRole, QuadraticBoundaryHamiltonian, QuadraticVolumeHamiltonian,
                    ScientificComponentSet, SurfaceRole, ContactRole, TiledCheckerboardCPM,
                    UnorderedContactHamiltonian,
                    compile_scientific_state, first_shell_relation, moore_relation,
                    required_properties)

# Float32 spacing, as PottsToolkit lowers a Float32 model (src/authoring/lowering.jl:473-484):
# relation weights are Float32, so no energy term is promoted to Float64 by them.
F32_SPACING = {2: (np.float32(1.0),) * 2, 3: (np.float32(1.0),) * 3}

J_TWO_TYPES = np.array([[0, 16, 16], [16, 2, 11], [16, 11, 14]], dtype=np.float32)


def make_algorithm(name: str, temperature):
    """CheckerboardSweepCPM, LotteryCPM or TiledCheckerboardCPM (reference defaults) by name."""
    if name == "checkerboard":
        return CheckerboardSweepCPM(temperature)
    if name == "lottery":
        return LotteryCPM(temperature)
    if name == "tiled":
        return TiledCheckerboardCPM(temperature)
    raise ValueError(f"unknown algorithm {name!r}")


def _schema(*components):
    schema = {}
    for c in components:
        schema.update(required_properties(c))
    return schema


def square_cells_2d(L: int, side: int, corners, ntypes: int = 2):
    """Owner words (column-major) and cell types for squares at the given corners."""
    owners = np.full((L, L), -1, dtype=np.int64)  # indexed [x, y]
    types = {}
    for k, (x0, y0) in enumerate(corners, start=1):
        owners[x0:x0 + side, y0:y0 + side] = k
        types[k] = 2 + (k - 1) % ntypes
    return owners.reshape(-1, order="F"), types


def plain_2d(L: int, corners, side: int = 10, *, temperature=np.float32(10.0), seed: int = 1234,
             lambda_v=np.float32(1.0), target_v=np.float32(100.0), lambda_s=None, target_s=40,
             algorithm="checkerboard", J=J_TWO_TYPES, ntypes: int = 2, contact="moore"):
    owners, types = square_cells_2d(L, side, corners, ntypes)
    ncells = len(corners)
    sp = F32_SPACING[2]
    domain = CartesianDomain((L, L), spacing=sp)
    proposal = first_shell_relation(ProposalRole, 2, spacing=sp)
    surface = first_shell_relation(SurfaceRole, 2, spacing=sp)
    contact_rel = (moore_relation(ContactRole, 2, spacing=sp) if contact == "moore"
                   else first_shell_relation(ContactRole, 2, spacing=sp))
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    contact_h = UnorderedContactHamiltonian(np.asarray(J, dtype=np.float32),
                                            MediumTypeTable({1: 1}), contact_rel)
    energies = [volume, contact_h]
    boundary = None
    if lambda_s is not None:
        boundary = QuadraticBoundaryHamiltonian(BoundaryEdgeCount(), surface, number_type=np.float32)
        energies.append(boundary)
    state = LogicalPottsState(owners, max(ncells, 1), cell_types=types, medium_domains=(1,),
                              property_schema=_schema(*energies), dims=(L, L))
    state.property_values("target_volume")[:ncells] = target_v
    state.property_values("volume_strength")[:ncells] = lambda_v
    if boundary is not None:
        state.property_values("target_boundary")[:ncells] = target_s
        state.property_values("boundary_strength")[:ncells] = lambda_s
    tracker = BoundaryMeasureTracker(BoundaryEdgeCount(), surface)
    compiled = compile_scientific_state(state, domain, tracker)
    alg = make_algorithm(algorithm, temperature)
    return compiled, proposal, ScientificComponentSet(energies=tuple(energies)), alg, seed


def config1_plain(algorithm="checkerboard", seed: int = 1234):
    """200x200, 100 cells of 10x10 at (5+20i, 5+20j), volume + Moore contact, T=10."""
    corners = [(5 + 20 * i, 5 + 20 * j) for j in range(10) for i in range(10)]
    return plain_2d(200, corners, algorithm=algorithm, seed=seed)


def grid_corners(L: int, side: int, count: int, pitch: int = 10):
    per_row = L // pitch
    out = []
    for k in range(count):
        row, col = divmod(k, per_row)
        out.append((pitch * col, pitch * row))
    return out


def config2_large(algorithm="checkerboard", L: int = 4096, cells: int = 163840, seed: int = 1234):
    """4096x4096, 163,840 cells of 10x10 on a pitch-10 slot grid filled row-major;
    volume + Moore contact + boundary (lambda_S = 1, S* = 40), T = 10."""
    corners = grid_corners(L, 10, cells)
    return plain_2d(L, corners, algorithm=algorithm, seed=seed, lambda_s=np.float32(1.0))


def scaled_2d(L: int, algorithm="checkerboard", seed: int = 1234, boundary: bool = True):
    """The config-2 layout rule on an L x L lattice (cells fill the slot grid)."""
    per_row = L // 10
    cells = per_row * per_row
    corners = grid_corners(L, 10, cells)
    return plain_2d(L, corners, algorithm=algorithm, seed=seed,
                    lambda_s=np.float32(1.0) if boundary else None)


def contact_matrix_3d(seed: int = 20261203) -> np.ndarray:
    """4x4 symmetric Float32 J with entries U[2, 20] from a fixed PCG64 stream."""
    rng = np.random.default_rng(seed)
    upper = rng.uniform(2.0, 20.0, size=(4, 4)).astype(np.float32)
    J = np.triu(upper) + np.triu(upper, 1).T
    return J.astype(np.float32)


def cubes_3d(L: int = 256, side: int = 8, *, temperature=np.float32(20.0), seed: int = 1234,
             algorithm="checkerboard", J=None):
    """L^3 periodic lattice tiled exactly by side^3 cubes (no medium sites), 3 types."""
    n_per = L // side
    idx = np.arange(L)
    cx, cy, cz = np.meshgrid(idx // side, idx // side, idx // side, indexing="ij")
    cell = 1 + cx + n_per * (cy + n_per * cz)
    owners = cell.reshape(-1, order="F").astype(np.int64)
    ncells = n_per ** 3
    types = {k: 2 + (k - 1) % 3 for k in range(1, ncells + 1)}
    sp = F32_SPACING[3]
    domain = CartesianDomain((L, L, L), spacing=sp)
    proposal = first_shell_relation(ProposalRole, 3, spacing=sp)
    surface = first_shell_relation(SurfaceRole, 3, spacing=sp)
    contact_rel = moore_relation(ContactRole, 3, spacing=sp)
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    J = contact_matrix_3d() if J is None else J
    contact_h = UnorderedContactHamiltonian(np.asarray(J, np.float32), MediumTypeTable({1: 1}),
                                            contact_rel)
    boundary = QuadraticBoundaryHamiltonian(BoundaryEdgeCount(), surface, number_type=np.float32)
    energies = (volume, contact_h, boundary)
    state = LogicalPottsState(owners, ncells, cell_types=types, medium_domains=(1,),
                              property_schema=_schema(*energies), dims=(L, L, L))
    state.property_values("target_volume")[:] = np.float32(side ** 3)
    state.property_values("volume_strength")[:] = np.float32(1.5)
    state.property_values("target_boundary")[:] = 6 * side * side
    state.property_values("boundary_strength")[:] = np.float32(0.5)
    compiled = compile_scientific_state(state, domain, BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    alg = make_algorithm(algorithm, temperature)
    return compiled, proposal, ScientificComponentSet(energies=energies), alg, seed


def block_cells(L: int, ndim: int, side: int, pitch: int, *, ntypes: int = 2, J=J_TWO_TYPES,
                temperature=np.float32(10.0), seed: int = 1234, lambda_v=np.float32(1.0),
                algorithm="checkerboard"):
    """L^ndim periodic lattice with side^ndim blocks on a pitch^ndim slot grid (every whole
    slot filled, slot k in first-axis-fastest order is cell k+1 of type 2 + k mod ntypes),
    medium id 1 (type 1) elsewhere; volume (lambda_v, V* = side^ndim) + Moore contact,
    von Neumann proposal, no boundary constraint."""
    per = L // pitch
    idx = np.arange(L)
    slot = idx // pitch
    inside = (idx % pitch) < side
    inside &= slot < per
    grids = np.meshgrid(*([slot] * ndim), indexing="ij")
    masks = np.meshgrid(*([inside] * ndim), indexing="ij")
    cell = np.zeros([L] * ndim, dtype=np.int64)
    stride = 1
    for g in grids:
        cell += g * stride
        stride *= per
    cell += 1
    occupied = np.logical_and.reduce(masks)
    owners = np.where(occupied, cell, -1).reshape(-1, order="F")
    ncells = per ** ndim
    types = {k: 2 + (k - 1) % ntypes for k in range(1, ncells + 1)}
    dims = (L,) * ndim
    sp = F32_SPACING[ndim]
    domain = CartesianDomain(dims, spacing=sp)
    proposal = first_shell_relation(ProposalRole, ndim, spacing=sp)
    surface = first_shell_relation(SurfaceRole, ndim, spacing=sp)
    contact_h = UnorderedContactHamiltonian(np.asarray(J, np.float32), MediumTypeTable({1: 1}),
                                            moore_relation(ContactRole, ndim, spacing=sp))
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    energies = (volume, contact_h)
    state = LogicalPottsState(owners, max(ncells, 1), cell_types=types, medium_domains=(1,),
                              property_schema=_schema(*energies), dims=dims)
    state.property_values("target_volume")[:ncells] = np.float32(side ** ndim)
    state.property_values("volume_strength")[:ncells] = lambda_v
    compiled = compile_scientific_state(state, domain, BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    alg = make_algorithm(algorithm, temperature)
    return compiled, proposal, ScientificComponentSet(energies=energies), alg, seed


def integer_contact_matrix(rng: np.random.Generator, ntypes: int) -> np.ndarray:
    """Symmetric (1 + ntypes)^2 Float32 J with integer entries uniform in {0, ..., 20}."""
    upper = rng.integers(0, 21, size=(ntypes + 1, ntypes + 1)).astype(np.float32)
    return (np.triu(upper) + np.triu(upper, 1).T).astype(np.float32)


ENSEMBLE_LAYOUT_SEED = 20261204


def ensemble_replica(r: int, master_seed: int = 1234, algorithm="checkerboard"):
    """Replica r (0-based) of the batched ensemble: side L = round(exp(U[ln 64, ln 1024])),
    10x10 cells on the pitch-10 slot grid, 2 types, T ~ U[2, 30], integer J ~ U{0..20}
    symmetric, lambda_V = 1; seed = ensemble_seed(master_seed, r + 1)."""
    from .native import ensemble_seed
    rng = np.random.default_rng([ENSEMBLE_LAYOUT_SEED, r])
    L = int(round(float(np.exp(rng.uniform(np.log(64.0), np.log(1024.0))))))
    T = np.float32(rng.uniform(2.0, 30.0))
    J = integer_contact_matrix(rng, 2)
    return block_cells(L, 2, 10, 10, ntypes=2, J=J, temperature=T,
                       seed=ensemble_seed(master_seed, r + 1), algorithm=algorithm)


SERVING_SEED = 20261205


class ServingJob(NamedTuple):
    index: int
    arrival_s: float       # seconds after the stream starts
    ndim: int
    side: int
    ntypes: int
    mcs: int
    seed: int


def serving_jobs(count: int, rate_per_s: float = 200.0, seed: int = SERVING_SEED,
                 min_2d: int = 128, max_2d: int = 2048, min_3d: int = 32, max_3d: int = 128):
    """The first `count` jobs of the served stream: Poisson arrivals at `rate_per_s`, 80 %
    2D squares of side log-uniform [min_2d, max_2d], 20 % 3D cubes of side log-uniform
    [min_3d, max_3d], 2-3 types, 100-2000 MCS each."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for k in range(count):
        t += float(rng.exponential(1.0 / rate_per_s))
        three_d = rng.uniform() < 0.2
        lo, hi = (min_3d, max_3d) if three_d else (min_2d, max_2d)
        side = int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))
        jobs.append(ServingJob(k, t, 3 if three_d else 2, side, int(rng.integers(2, 4)),
                               int(rng.integers(100, 2001)), int(rng.integers(0, 2 ** 63))))
    return jobs


def serving_model(job: "ServingJob", algorithm="checkerboard"):
    """Model of one served job: cells of 10^2 (2D) or 8^3 (3D) on slot grids of the same
    pitch, integer J ~ U{0..20} symmetric drawn from the job seed, lambda_V = 1, T = 10."""
    rng = np.random.default_rng([SERVING_SEED, job.index])
    J = integer_contact_matrix(rng, job.ntypes)
    side = 10 if job.ndim == 2 else 8
    return block_cells(job.side, job.ndim, side, side, ntypes=job.ntypes, J=J,
                       temperature=np.float32(10.0), seed=job.seed, algorithm=algorithm)


def single_cell_migration(algorithm="checkerboard", seed: int = 0, L: int = 64):
    """The reference's realistic single-cell-migration workload
    (integration/transition/RealisticScaleRunner.jl:126-148 with the manifest's model,
    design/audits/phase-13-realistic-workload-manifest.toml): an L x L periodic lattice of
    medium id 1 (type 1) holding cell 1 (type 2) as a centred 5 x 5 square (sites
    fld(L - 5, 2) + 1 ... + 5 of each axis, 1-based); as PottsToolkit lowers it
    (src/authoring/lowering.jl:468-505): first-shell von Neumann proposal, contact and surface
    relations, energies (volume, contact) in Float32 with target volume 25 and strength 1,
    contact J (medium-medium 0, medium-cell 12, cell-cell 4), a BoundaryEdgeCount tracker,
    capacity 1, T = 5. `seed` is the replica's dynamics seed."""
    side = 5
    start = (L - side) // 2
    owners = np.full((L, L), -1, dtype=np.int64)
    owners[start:start + side, start:start + side] = 1
    f32 = np.float32
    spacing = (f32(1.0), f32(1.0))
    proposal = first_shell_relation(ProposalRole, 2, spacing=spacing)
    surface = first_shell_relation(SurfaceRole, 2, spacing=spacing)
    contact_rel = first_shell_relation(ContactRole, 2, spacing=spacing)
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    contact = UnorderedContactHamiltonian(np.array([[0, 12], [12, 4]], dtype=np.float32),
                                          MediumTypeTable({1: 1}), contact_rel)
    energies = (volume, contact)
    state = LogicalPottsState(owners, 1, cell_types={1: 2}, medium_domains=(1,),
                              property_schema=_schema(*energies))
    state.property_values("target_volume")[:] = f32(25.0)
    state.property_values("volume_strength")[:] = f32(1.0)
    compiled = compile_scientific_state(state, CartesianDomain((L, L), spacing=spacing),
                                        BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    alg = make_algorithm(algorithm, f32(5.0))
    return compiled, proposal, ScientificComponentSet(energies=energies), alg, seed
