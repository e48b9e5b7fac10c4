"""TiledCheckerboardCPM on the MI355X engine against the oracle, bit for bit
(lib/CorePotts/test/test_tiled_checkerboard.jl): the reference's 6 x 6 replay fixture, tissues
in 2D and 3D (periodic and closed, default and explicit tiles, switching intervals that split
a tile's proposals into batches), the colour order of every MCS, a config-2-rule lattice
whose 4 x 4 tiles run many colours' worth of tiles at once, graph-captured steps, a batch of
tiled replicas, and exact continuation through a checkpoint. The oracle's additive
reconciliation (what the device does) is pinned to its recount reconciliation by the oracle
tests."""
import contextlib

import numpy as np
import pytest

from corepotts import ReplicaBatch, TiledCheckerboardCPM, init_scientific, set_tiled_staging, workloads
from corepotts import checkpoint as ck
from corepotts.spec import ALG_TILED
from helpers import oracle_for
from test_oracle_tiled import _tissue, reference_fixture

pytestmark = pytest.mark.gpu

FIELDS = ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures", "active", "cell_types",
          "target_volume", "volume_strength", "target_boundary", "boundary_strength")


def _same(g, o):
    for k in FIELDS:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)


@contextlib.contextmanager
def staging(enable):
    """The staged (LDS) or the global-memory tiled kernel for models built inside."""
    set_tiled_staging(enable)
    try:
        yield
    finally:
        set_tiled_staging(True)


def test_gpu_reference_fixture_replay():
    # test_tiled_checkerboard.jl:39-90: tile (2, 2), switching interval 2, 3 MCS
    compiled, proposal, components = reference_fixture()
    algorithm = TiledCheckerboardCPM(np.float32(5.0), tile_size=(2, 2), switching_interval=2)
    print("tiled reference fixture", flush=True)
    gpu = init_scientific(compiled, proposal, components, algorithm, seed=0x12C5)
    sim, _ = oracle_for(compiled, proposal, components, algorithm, 0x12C5)
    for _ in range(3):
        gpu.step(1)
        sim.step(1)
        _same(gpu.download(), sim.store())
        assert tuple(gpu.current_mcs_report()) == sim.report()
        count, order = gpu.schedule()
        np.testing.assert_array_equal(order, sim.last_order())
    assert gpu.current_mcs_report()[1] == count * 2
    gpu.close()


CASES = [
    ((24, 20), 1, "moore", True, True, None, None),
    ((23, 17), 2, "vn", True, False, (4, 3), 5),
    ((30, 30), 3, "moore", False, True, (3, 5), None),
    ((10, 9, 8), 4, "moore", True, True, None, 3),
    ((9, 8, 7), 5, "vn", True, False, (2, 3, 2), None),
    ((11, 11), 6, "vn", True, True, (3, 3), 4),
]


@pytest.mark.parametrize("staged", [True, False], ids=["staged", "global"])
@pytest.mark.parametrize("case", CASES, ids=["2d-default", "2d-closed-4x3", "2d-3x5", "3d-default", "3d-closed",
                                             "2d-seam-3x3"])
def test_gpu_tiled_tissue_parity(case, staged):
    dims, seed, contact, boundary, periodic, tile, interval = case
    compiled, proposal, components = _tissue(dims, seed, None, contact, boundary, periodic)
    algorithm = TiledCheckerboardCPM(np.float32(12.0), tile_size=tile, switching_interval=interval)
    print(f"tiled tissue {dims} tile {tile} interval {interval} staged {staged}", flush=True)
    with staging(staged):
        gpu = init_scientific(compiled, proposal, components, algorithm, seed=1000 + seed)
    sim, _ = oracle_for(compiled, proposal, components, algorithm, 1000 + seed)
    accepted = 0
    for _ in range(6):
        gpu.step(1)
        sim.step(1)
        _same(gpu.download(), sim.store())
        report = tuple(gpu.current_mcs_report())
        assert report == sim.report()
        accepted += report[11]
    assert accepted > 0
    gpu.close()


@pytest.mark.parametrize("graph,staged", [(False, True), (True, True), (False, False)],
                         ids=["eager", "graph", "global"])
def test_gpu_tiled_config2_rule_lattice(graph, staged):
    # the config-2 layout rule (10 x 10 cells on a pitch-10 grid, volume + Moore contact +
    # boundary) at 200 x 200 with the default 4 x 4 tiles: 2500 tiles, 625 per colour
    model = workloads.scaled_2d(200, algorithm="tiled", seed=77)
    print(f"tiled config-2 rule 200x200 graph={graph} staged={staged}", flush=True)
    with staging(staged):
        gpu = init_scientific(*model[:4], seed=model[4])
    if graph:
        gpu.set_graph_chunk(2)
    sim, _ = oracle_for(*model)
    gpu.step(4)
    sim.step(4)
    _same(gpu.download(), sim.store())
    assert tuple(gpu.current_mcs_report()) == sim.report()
    gpu.close()


@pytest.mark.parametrize("staged", [True, False], ids=["staged", "global"])
@pytest.mark.parametrize("dims,count,periodic", [((40, 36), 30, True), ((33, 30), 60, False), ((12, 10, 14), 20, True)],
                         ids=["2d-periodic", "2d-closed", "3d-periodic"])
def test_gpu_tiled_obstacles(dims, count, periodic, staged):
    # obstacle sites: recipients from the tile site lists, obstacles never donate
    from test_colouring import _obstacle_domain
    compiled, proposal, components = _obstacle_domain(dims, count, 5 + count, periodic)
    algorithm = TiledCheckerboardCPM(np.float32(9.0))
    print(f"tiled obstacles {dims} staged {staged}", flush=True)
    with staging(staged):
        gpu = init_scientific(compiled, proposal, components, algorithm, seed=31)
    sim, _ = oracle_for(compiled, proposal, components, algorithm, 31)
    for _ in range(4):
        gpu.step(1)
        sim.step(1)
        _same(gpu.download(), sim.store())
        assert tuple(gpu.current_mcs_report()) == sim.report()
    gpu.close()


@pytest.mark.parametrize("staged", [True, False], ids=["staged", "global"])
def test_gpu_tiled_many_cells_per_tile(staged):
    # one cell per site at a high temperature: a tile's 16 proposals touch up to 32 cells,
    # more than the kernels keep in LDS (the rest go to the global spill area); cells that
    # lose their only site retire at the reconciliation
    from corepotts import (BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain, LogicalPottsState,
                           ProposalRole, QuadraticVolumeHamiltonian, ScientificComponentSet, SurfaceRole,
                           compile_scientific_state, first_shell_relation, required_properties)
    dims = (24, 20)
    n = dims[0] * dims[1]
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    owners = np.arange(1, n + 1, dtype=np.int64).reshape(dims, order="F")
    state = LogicalPottsState(owners, n, cell_types={c: 1 for c in range(1, n + 1)}, medium_domains=(1,),
                              property_schema=required_properties(volume))
    state.property_values("target_volume")[:] = 2.0
    state.property_values("volume_strength")[:] = 0.5
    surface = first_shell_relation(SurfaceRole, 2)
    compiled = compile_scientific_state(state, CartesianDomain(dims), BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    components = ScientificComponentSet(energies=(volume,))
    algorithm = TiledCheckerboardCPM(np.float32(50.0))
    proposal = first_shell_relation(ProposalRole, 2)
    print(f"tiled one-site cells staged {staged}", flush=True)
    with staging(staged):
        gpu = init_scientific(compiled, proposal, components, algorithm, seed=77)
    sim, _ = oracle_for(compiled, proposal, components, algorithm, 77)
    for _ in range(4):
        gpu.step(1)
        sim.step(1)
        _same(gpu.download(), sim.store())
        assert tuple(gpu.current_mcs_report()) == sim.report()
    assert sim.report()[11] > 100
    gpu.close()


def test_gpu_tiled_batch():
    print("tiled batch", flush=True)
    batch = ReplicaBatch(ALG_TILED)
    models = []
    for k, case in enumerate(CASES[:3]):
        dims, seed, contact, boundary, periodic, tile, interval = case
        compiled, proposal, components = _tissue(dims, seed, None, contact, boundary, periodic)
        models.append((compiled, proposal, components,
                       TiledCheckerboardCPM(np.float32(12.0), tile_size=tile, switching_interval=interval), 50 + k))
    slots = [batch.add_model(*m, target_mcs=3 + k) for k, m in enumerate(models)]
    batch.step(6)
    for k, (slot, m) in enumerate(zip(slots, models)):
        sim, _ = oracle_for(*m)
        sim.step(3 + k)
        _same(batch.download(slot), sim.store())
        assert tuple(batch.report(slot)) == sim.report()
    batch.close()


def test_gpu_tiled_checkpoint_continuation():
    model = workloads.scaled_2d(60, algorithm="tiled", seed=0xBEEF)
    print("tiled checkpoint", flush=True)
    gpu = init_scientific(*model[:4], seed=model[4])
    gpu.step(3)
    cp = ck.capture_checkpoint(gpu)
    gpu.step(4)
    resumed = ck.restore_checkpoint(cp, gpu)
    resumed.step(4)
    _same(resumed.download(), gpu.download())
    assert resumed.current_mcs_report() == gpu.current_mcs_report()
    resumed.close()
    gpu.close()
