"""TiledCheckerboardCPM in the oracle, following lib/CorePotts/test/test_tiled_checkerboard.jl.

The reference checks its device path against its host reference (same lattice, same report,
trackers that match a recount). Here the oracle runs both reconciliations, snapshot plus
additive per-tile deltas (the device path) and a recount from ownership after every
subround (the host reference), and they must agree bit for bit; the layout, the halo
colouring (against the reference's all-pairs greedy scan, restated below) and the colour
order are checked as the reference's tests check them."""
import itertools

import numpy as np
import pytest

from corepotts import (AxisBoundary, BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain,
                       CellOwner, ClosedBoundary, ContactRole, LogicalPottsState, MediumOwner,
                       MediumTypeTable, PeriodicBoundary, ProposalRole, QuadraticBoundaryHamiltonian,
                       QuadraticVolumeHamiltonian, ScientificComponentSet, SurfaceRole,
                       TiledCheckerboardCPM, UnorderedContactHamiltonian, build_spec,
                       compile_scientific_state, first_shell_relation, moore_relation,
                       required_properties, state_arrays, workloads)
from oracle import bridge


def reference_fixture(dims=(6, 6), boundaries=None):
    """test_tiled_checkerboard.jl:7-37: two 2^d cells in opposite corners of a medium."""
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    owners = np.full(dims, MediumOwner(1), dtype=np.int64)
    corner = tuple(slice(0, 2) for _ in dims)
    far = tuple(slice(d - 2, d) for d in dims)
    owners[corner] = CellOwner(1)
    owners[far] = CellOwner(2)
    state = LogicalPottsState(owners, 2, cell_types={1: 2, 2: 2}, medium_domains=(1,),
                              property_schema=required_properties(volume))
    state.property_values("target_volume")[:] = 5.0
    state.property_values("volume_strength")[:] = 1.25
    ndim = len(dims)
    domain = CartesianDomain(dims) if boundaries is None else CartesianDomain(dims, boundaries=boundaries)
    contact = UnorderedContactHamiltonian(np.array([[0, 3], [3, 1]], np.float32), MediumTypeTable({1: 1}),
                                          first_shell_relation(ContactRole, ndim))
    tracker = BoundaryMeasureTracker(BoundaryEdgeCount(), first_shell_relation(SurfaceRole, ndim))
    compiled = compile_scientific_state(state, domain, tracker)
    return compiled, first_shell_relation(ProposalRole, ndim), ScientificComponentSet(energies=(volume, contact))


def _sims(compiled, proposal, components, algorithm, seed):
    spec = build_spec(compiled, proposal, components, algorithm, seed)
    additive = bridge.OracleSim(spec, state_arrays(compiled, spec))
    recount = bridge.OracleSim(spec, state_arrays(compiled, spec))
    recount.set_tiled_reconcile(True)
    return spec, additive, recount


def _assert_same(a, b):
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def _tissue(dims, seed, algorithm, contact="moore", boundary=True, periodic=True):
    rng = np.random.default_rng(seed)
    owners = np.full(dims, MediumOwner(1), dtype=np.int64)
    cid = 0
    side = 3
    for origin in itertools.product(*[range(1, d - side, side + 1) for d in dims]):
        if rng.random() < 0.8 or cid == 0:
            cid += 1
            owners[tuple(slice(o, o + side) for o in origin)] = CellOwner(cid)
    ndim = len(dims)
    surface = first_shell_relation(SurfaceRole, ndim)
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    rel = moore_relation(ContactRole, ndim) if contact == "moore" else first_shell_relation(ContactRole, ndim)
    contact_h = UnorderedContactHamiltonian(np.array([[0, 5, 7], [5, 2, 6], [7, 6, 3]], np.float32),
                                            MediumTypeTable({1: 1}), rel)
    energies = [volume, contact_h]
    schema = dict(required_properties(volume))
    if boundary:
        bnd = QuadraticBoundaryHamiltonian(BoundaryEdgeCount(), surface, number_type=np.float32)
        energies.append(bnd)
        schema.update(required_properties(bnd))
    state = LogicalPottsState(owners, cid + 3, cell_types={c: 2 + c % 2 for c in range(1, cid + 1)},
                              medium_domains=(1,), property_schema=schema)
    state.property_values("target_volume")[:cid] = float(side ** ndim)
    state.property_values("volume_strength")[:cid] = 1.0
    if boundary:
        state.property_values("target_boundary")[:cid] = 2 * ndim * side ** (ndim - 1)
        state.property_values("boundary_strength")[:cid] = 0.5
    b = PeriodicBoundary() if periodic else ClosedBoundary()
    domain = CartesianDomain(dims, boundaries=tuple(AxisBoundary(b) for _ in dims))
    compiled = compile_scientific_state(state, domain, BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    return compiled, first_shell_relation(ProposalRole, ndim), ScientificComponentSet(energies=tuple(energies))


def test_reference_fixture_replay_accounting_and_reconciliations():
    # test_tiled_checkerboard.jl:39-90: tile (2, 2), switching interval 2, 3 MCS
    compiled, proposal, components = reference_fixture()
    algorithm = TiledCheckerboardCPM(np.float32(5.0), tile_size=(2, 2), switching_interval=2)
    spec, first, recount = _sims(compiled, proposal, components, algorithm, 0x12C5)
    second = bridge.OracleSim(spec, state_arrays(compiled, spec))
    for sim in (first, second, recount):
        sim.step(3)
    _assert_same(first.store(), second.store())
    _assert_same(first.store(), recount.store())
    report = first.report()
    assert report == second.report() == recount.report()
    mutable = int(compiled.domain.mutable_site_count())
    assert report[0] == 3
    assert report[2] == report[3] == mutable
    assert report[3] == report[4] + report[5] + report[6] + report[7]
    assert report[4] == report[10] + report[11]
    grid, reach, halo, colors, count = first.tiled_layout()
    assert report[1] == count * 2          # two batches of two proposals per 2 x 2 tile
    assert first.tracker_errors() == 0


@pytest.mark.parametrize("case", [
    ((24, 20), 1, "moore", True, True, None, None),
    ((23, 17), 2, "vn", True, False, (4, 3), 5),
    ((30, 30), 3, "moore", False, True, (3, 5), None),
    ((10, 9, 8), 4, "moore", True, True, None, 3),
    ((9, 8, 7), 5, "vn", True, False, (2, 3, 2), None),
], ids=["2d-default", "2d-closed-4x3", "2d-3x5", "3d-default", "3d-closed"])
def test_additive_reconciliation_equals_recount(case):
    dims, seed, contact, boundary, periodic, tile, interval = case
    compiled, proposal, components = _tissue(dims, seed, None, contact, boundary, periodic)
    algorithm = TiledCheckerboardCPM(np.float32(12.0), tile_size=tile, switching_interval=interval)
    spec, additive, recount = _sims(compiled, proposal, components, algorithm, 1000 + seed)
    accepted = 0
    for _ in range(6):
        additive.step(1)
        recount.step(1)
        _assert_same(additive.store(), recount.store())
        assert additive.report() == recount.report()
        assert additive.tracker_errors() == 0
        r = additive.report()
        assert r[3] == r[4] + r[5] + r[6] + r[7]
        assert r[8] == 0 and r[4] == r[10] + r[11]
        accepted += r[11]
    assert accepted > 0


def _reference_greedy(grid, periodic, reach):
    """_tiled_greedy_coloring (tiled_checkerboard.jl:131-158): all prior tiles, linear order."""
    coords = list(itertools.product(*[range(g) for g in reversed(grid)]))
    coords = [tuple(reversed(c)) for c in coords]        # first axis fastest

    def conflict(a, b):
        for d in range(len(grid)):
            direct = abs(a[d] - b[d])
            sep = min(direct, grid[d] - direct) if periodic[d] else direct
            if sep > reach[d]:
                return False
        return True

    colors = []
    for t, c in enumerate(coords):
        used = {colors[p] for p in range(t) if conflict(c, coords[p])}
        col = 1
        while col in used:
            col += 1
        colors.append(col)
    return np.array(colors, np.uint16)


@pytest.mark.parametrize("dims,tile,periodic", [
    ((9, 7), (4, 3), False), ((12, 12), (4, 4), True), ((11, 11), (3, 3), True), ((10, 6), (1, 1), True),
    ((7, 5), (2, 2), True), ((8, 6, 6), (2, 2, 2), True), ((9, 7, 5), (4, 3, 2), False),
])
def test_halo_colouring_matches_the_reference_greedy_scan(dims, tile, periodic):
    # test_tiled_checkerboard.jl:275-300
    b = PeriodicBoundary() if periodic else ClosedBoundary()
    compiled, proposal, components = reference_fixture(dims, tuple(AxisBoundary(b) for _ in dims))
    spec = build_spec(compiled, proposal, components, TiledCheckerboardCPM(tile_size=tile), 1)
    sim = bridge.OracleSim(spec, state_arrays(compiled, spec))
    grid, reach, halo, colors, count = sim.tiled_layout()
    nd = len(dims)
    assert grid[:nd] == tuple(-(-d // t) for d, t in zip(dims, tile))
    assert halo == 1 and reach[:nd] == tuple(-(-1 // t) for t in tile)
    np.testing.assert_array_equal(colors, _reference_greedy(grid[:nd], (periodic,) * nd, reach[:nd]))
    assert count == colors.max()
    if dims == (12, 12):
        assert count >= 3


def test_colour_order_is_a_replayable_permutation():
    # test_tiled_checkerboard.jl:302-318
    compiled, proposal, components = reference_fixture((11, 11), (AxisBoundary(PeriodicBoundary()),) * 2)
    algorithm = TiledCheckerboardCPM(tile_size=(3, 3))
    spec = build_spec(compiled, proposal, components, algorithm, 0x125)
    a = bridge.OracleSim(spec, state_arrays(compiled, spec))
    b = bridge.OracleSim(spec, state_arrays(compiled, spec))
    orders = set()
    for _ in range(12):
        a.step(1)
        b.step(1)
        oa = tuple(int(v) for v in a.last_order())
        assert oa == tuple(int(v) for v in b.last_order())
        assert sorted(oa) == list(range(1, len(oa) + 1))
        orders.add(oa)
    assert len(orders) > 1


def test_tiled_configuration_validation():
    # test_tiled_checkerboard.jl:230-250
    algorithm = TiledCheckerboardCPM(np.float32(6.0), tile_size=(4, 3), switching_interval=7, shared_memory="required")
    assert algorithm.temperature == np.float32(6.0) and algorithm.tile_size == (4, 3)
    assert algorithm.switching_interval == 7 and algorithm.shared_memory == "required"
    assert TiledCheckerboardCPM().shared_memory == "auto"
    for bad in (dict(temperature=-1.0), dict(tile_size=(4,)), dict(tile_size=(4, 0)), dict(switching_interval=0),
                dict(shared_memory="fallback")):
        with pytest.raises(ValueError):
            TiledCheckerboardCPM(**bad)
    compiled, proposal, components = reference_fixture()
    with pytest.raises(ValueError):
        build_spec(compiled, proposal, components, TiledCheckerboardCPM(tile_size=(2, 2, 2)), 1)
