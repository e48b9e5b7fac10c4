"""Every compiled tile shape gives the oracle's results bit for bit (the shape only changes
launch geometry): 2D checkerboard and lottery on odd and even sides, and 3D float-J cubes."""
import numpy as np
import pytest

from corepotts import init_scientific, set_row_kernels, set_tile_shape, tile_shapes, workloads
from corepotts.spec import ALG_CHECKERBOARD, ALG_LOTTERY
from helpers import oracle_for

pytestmark = pytest.mark.gpu

FIELDS = ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures", "active",
          "cell_types")


def _parity(model, mcs, rows=None):
    compiled, proposal, components, alg, seed = model
    gpu = init_scientific(compiled, proposal, components, alg, seed=seed)
    if rows is not None:
        assert bool(gpu.kernel_config() & 32) == rows
    sim, _ = oracle_for(compiled, proposal, components, alg, seed)
    gpu.step(mcs)
    sim.step(mcs)
    g, o = gpu.download(), sim.store()
    for f in FIELDS:
        np.testing.assert_array_equal(g[f], o[f], err_msg=f)
    assert tuple(gpu.current_mcs_report()) == sim.report()
    gpu.close()


@pytest.mark.parametrize("shape", [s for s, _ in tile_shapes(2, 1)])
@pytest.mark.parametrize("alg", ["checkerboard", "lottery"])
def test_2d_shapes(gpu_lib, shape, alg):
    alg_id = ALG_CHECKERBOARD if alg == "checkerboard" else ALG_LOTTERY
    set_tile_shape(alg_id, 2, 1, shape)
    try:
        for L in (97, 130):
            _parity(workloads.scaled_2d(L, algorithm=alg, seed=4242 + L), 4)
    finally:
        set_tile_shape(alg_id, 2, 1, -1)


def _rectangle(X, Y, seed):
    """X x Y periodic lattice of 10 x 10 cells (volume, Moore contact, boundary), medium gaps."""
    from corepotts import (BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain, CheckerboardSweepCPM,
                           ContactRole, LogicalPottsState, MediumTypeTable, ProposalRole,
                           QuadraticBoundaryHamiltonian, QuadraticVolumeHamiltonian, ScientificComponentSet,
                           SurfaceRole, UnorderedContactHamiltonian, compile_scientific_state,
                           first_shell_relation, moore_relation, required_properties)
    owners = np.full((X, Y), -1, dtype=np.int64)
    types, cid = {}, 0
    for y0 in range(0, Y - 9, 12):
        for x0 in range(0, X - 9, 12):
            cid += 1
            owners[x0:x0 + 10, y0:y0 + 10] = cid
            types[cid] = 2 + cid % 2
    surface = first_shell_relation(SurfaceRole, 2)
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    contact = UnorderedContactHamiltonian(np.asarray(workloads.J_TWO_TYPES, np.float32), MediumTypeTable({1: 1}),
                                          moore_relation(ContactRole, 2))
    boundary = QuadraticBoundaryHamiltonian(BoundaryEdgeCount(), surface, number_type=np.float32)
    schema = {**required_properties(volume), **required_properties(boundary)}
    state = LogicalPottsState(owners, cid, cell_types=types, medium_domains=(1,), property_schema=schema)
    state.property_values("target_volume")[:] = np.float32(100)
    state.property_values("volume_strength")[:] = np.float32(1)
    state.property_values("target_boundary")[:] = 40
    state.property_values("boundary_strength")[:] = np.float32(1)
    compiled = compile_scientific_state(state, CartesianDomain((X, Y)), BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    return (compiled, first_shell_relation(ProposalRole, 2), ScientificComponentSet(energies=(volume, contact, boundary)),
            CheckerboardSweepCPM(np.float32(10.0)), seed)


ROW_SHAPES = [s for s, t in tile_shapes(2, 1) if t[0] * t[1] >= 1024]


@pytest.mark.parametrize("shape", ROW_SHAPES)
@pytest.mark.parametrize("rows", [True, False], ids=["rows", "tiles"])
def test_2d_row_kernels(gpu_lib, shape, rows):
    """Lattices whose sides are whole tiles of every shape: the row propose kernel (periodic,
    von Neumann proposal, Moore contact: a 2 x 2 colour pattern) against the oracle, and the
    tile kernel on the same models."""
    set_tile_shape(ALG_CHECKERBOARD, 2, 1, shape)
    set_row_kernels(rows)
    try:
        _parity(workloads.scaled_2d(256, seed=9001), 5, rows)
        _parity(_rectangle(512, 64, seed=9002), 4, rows)
    finally:
        set_tile_shape(ALG_CHECKERBOARD, 2, 1, -1)
        set_row_kernels(True)


@pytest.mark.parametrize("shape", ROW_SHAPES)
def test_2d_row_kernels_half_rows(gpu_lib, shape):
    """x sides of 2 mod 4: every other row starts 8 bytes into a 16-byte group and each row
    ends in a group of two sites, whose last site wraps to the row's first word (the row
    kernel's variant of 8-byte halves, kernel configuration bits 5 and 6)."""
    set_tile_shape(ALG_CHECKERBOARD, 2, 1, shape)
    try:
        for model in (workloads.scaled_2d(130, seed=9003), _rectangle(66, 40, seed=9004),
                      _rectangle(22, 26, seed=9005)):
            gpu = init_scientific(*model[:4], seed=model[4])
            assert gpu.kernel_config() & 96 == 96
            gpu.close()
            _parity(model, 5, True)
    finally:
        set_tile_shape(ALG_CHECKERBOARD, 2, 1, -1)


@pytest.mark.parametrize("shape", [s for s, _ in tile_shapes(3, 1)])
@pytest.mark.parametrize("alg", ["checkerboard", "lottery"])
def test_3d_shapes(gpu_lib, shape, alg):
    alg_id = ALG_CHECKERBOARD if alg == "checkerboard" else ALG_LOTTERY
    set_tile_shape(alg_id, 3, 1, shape)
    try:
        _parity(workloads.cubes_3d(24, side=6, algorithm=alg), 3)
        _parity(workloads.block_cells(21, 3, 5, 7, algorithm=alg, seed=77), 3)
    finally:
        set_tile_shape(alg_id, 3, 1, -1)


ROW3_SHAPES = [s for s, t in tile_shapes(3, 1) if s in (2, 8, 9)]


@pytest.mark.parametrize("shape", ROW3_SHAPES)
@pytest.mark.parametrize("rows", [True, False], ids=["rows", "tiles"])
def test_3d_row_kernels(gpu_lib, shape, rows):
    """Periodic cubes whose sides are whole tiles of every 3D shape: the 3D row propose kernel
    (von Neumann proposal, 26-neighbour contact: a 2 x 2 x 2 colour pattern, eight colours)
    against the oracle, and the tile kernel on the same models."""
    set_tile_shape(ALG_CHECKERBOARD, 3, 1, shape)
    set_row_kernels(rows)
    try:
        _parity(workloads.cubes_3d(64, side=8, seed=515), 3, rows)
        _parity(workloads.block_cells(64, 3, 6, 8, seed=516), 3, rows)
    finally:
        set_tile_shape(ALG_CHECKERBOARD, 3, 1, -1)
        set_row_kernels(True)


@pytest.mark.parametrize("shape", ROW_SHAPES)
def test_2d_row_kernels_partial_tiles(gpu_lib, shape):
    """Sides that are multiples of 4 but not of the tile: the tiles at the upper x and y edges
    of the row kernel hang over the lattice and mask their missing rows and 16-byte groups."""
    set_tile_shape(ALG_CHECKERBOARD, 2, 1, shape)
    try:
        _parity(workloads.scaled_2d(260, seed=9003), 4, True)
        _parity(_rectangle(516, 90, seed=9004), 4, True)
    finally:
        set_tile_shape(ALG_CHECKERBOARD, 2, 1, -1)


@pytest.mark.parametrize("shape", ROW3_SHAPES)
def test_3d_row_kernels_partial_tiles(gpu_lib, shape):
    """68^3 periodic cubes: every 3D row tile shape overhangs the lattice on at least one axis."""
    set_tile_shape(ALG_CHECKERBOARD, 3, 1, shape)
    try:
        _parity(workloads.cubes_3d(68, side=4, seed=517), 3, True)
        _parity(workloads.block_cells(68, 3, 6, 8, seed=518), 3, True)
    finally:
        set_tile_shape(ALG_CHECKERBOARD, 3, 1, -1)


def _site_cells(dims, seed):
    """Every site its own cell (volume target 2, Moore contact, boundary tracker) at a high
    temperature: a tile's attempts touch more distinct cells than the claim table holds, so
    most claims take the global-atomic path."""
    from corepotts import (BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain, CheckerboardSweepCPM,
                           ContactRole, LogicalPottsState, MediumTypeTable, ProposalRole,
                           QuadraticVolumeHamiltonian, ScientificComponentSet, SurfaceRole,
                           UnorderedContactHamiltonian, compile_scientific_state, first_shell_relation,
                           moore_relation, required_properties)
    ndim = len(dims)
    n = int(np.prod(dims))
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    contact = UnorderedContactHamiltonian(np.array([[0, 4, 4], [4, 2, 6], [4, 6, 3]], np.float32),
                                          MediumTypeTable({1: 1}), moore_relation(ContactRole, ndim))
    owners = np.arange(1, n + 1, dtype=np.int64).reshape(dims, order="F")
    state = LogicalPottsState(owners, n, cell_types={c: 2 + c % 2 for c in range(1, n + 1)}, medium_domains=(1,),
                              property_schema=required_properties(volume))
    state.property_values("target_volume")[:] = 2.0
    state.property_values("volume_strength")[:] = 0.5
    surface = first_shell_relation(SurfaceRole, ndim)
    compiled = compile_scientific_state(state, CartesianDomain(dims), BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    return (compiled, first_shell_relation(ProposalRole, ndim), ScientificComponentSet(energies=(volume, contact)),
            CheckerboardSweepCPM(np.float32(40.0)), seed)


@pytest.mark.parametrize("rows", [True, False], ids=["rows", "tiles"])
@pytest.mark.parametrize("dims", [(256, 32), (32, 16, 16)], ids=["2d", "3d"])
def test_claim_table_overflow(gpu_lib, rows, dims):
    """One cell per site: the propose kernels' per-tile claim tables overflow and fall back to
    global atomics for most cells, with results bit-exact against the oracle."""
    set_row_kernels(rows)
    try:
        _parity(_site_cells(dims, 4711), 3, rows)
    finally:
        set_row_kernels(True)
