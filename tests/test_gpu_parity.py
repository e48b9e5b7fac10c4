"""HIP engine versus the CPU oracle: bit-exact lattice, trackers and reports.

Integer-valued models (all configurations here except where noted) must agree exactly
after every compared MCS. Float32 models (the 3D random-J case) are compared the same way:
the energy change is evaluated in the reference's operation order on both sides with
contraction disabled, and acceptance uses a double-precision exp rounded once to float.
"""
import numpy as np
import pytest

from corepotts import (AxisBoundary, BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain,
                       CellOwner, CheckerboardSweepCPM, ClosedBoundary, ContactRole, FixedExterior,
                       LogicalPottsState, LotteryCPM, MediumOwner, MediumTypeTable, PeriodicBoundary,
                       ProposalRole, QuadraticBoundaryHamiltonian, QuadraticVolumeHamiltonian,
                       ScientificComponentSet, SurfaceRole, UnorderedContactHamiltonian,
                       compile_scientific_state, first_shell_relation, init_scientific,
                       moore_relation, required_properties, tracker_conformance_errors)
from corepotts import workloads
from helpers import load_phase13, oracle_for, phase13_fixture, scientific_fixture

pytestmark = pytest.mark.gpu

FIELDS = ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures", "active",
          "cell_types", "target_volume", "volume_strength", "target_boundary", "boundary_strength")


def _compare(gpu, sim, label):
    g = gpu.download()
    o = sim.store()
    for f in FIELDS:
        np.testing.assert_array_equal(g[f], o[f], err_msg=f"{label}: {f} differs")
    rep = gpu.current_mcs_report()
    assert rep is not None
    assert tuple(rep) == sim.report(), f"{label}: report differs"
    return rep


def run_parity(compiled, proposal, components, alg, seed, checkpoints):
    gpu = init_scientific(compiled, proposal, components, alg, seed=seed)
    sim, _ = oracle_for(compiled, proposal, components, alg, seed)
    done = 0
    reps = []
    for target in checkpoints:
        gpu.step(target - done)
        sim.step(target - done)
        done = target
        reps.append(_compare(gpu, sim, f"mcs {target}"))
    assert sim.tracker_errors() == 0
    assert not tracker_conformance_errors(gpu.logical_state())
    return gpu, sim, reps


def _partition(rep):
    assert rep.realized_proposals == (rep.dynamic_conflicts + rep.constraint_rejections +
                                      rep.acceptance_rejections + rep.accepted_copies)
    assert rep.activated_attempts == (rep.same_owner_no_ops + rep.boundary_no_ops +
                                      rep.immutable_recipient_no_ops + rep.dynamic_conflicts +
                                      rep.constraint_rejections + rep.acceptance_rejections +
                                      rep.accepted_copies)


@pytest.mark.parametrize("row_index", [0, 2, 4, 6])
def test_phase13_rows_per_seed(gpu_lib, row_index):
    data = load_phase13()
    row = data["rows"][row_index]
    assert row["algorithm"] == "checkerboard"
    compiled, proposal, components, alg, labels = phase13_fixture(row, data["model"], CheckerboardSweepCPM)
    gpu = init_scientific(compiled, proposal, components, alg, seed=0)
    sim, _ = oracle_for(compiled, proposal, components, alg, 0)
    pristine = sim.store()
    base = int(row["first_seed_hex"], 16)
    counts = np.zeros(len(labels) ** len(compiled.owners), dtype=np.int64)
    for r in range(512):
        gpu.upload(compiled)
        gpu.set_seed(base + r)
        gpu.mcs = 0
        gpu.step(1)
        g = gpu.download()
        sim.load(pristine)
        sim.set_seed(base + r)
        sim.mcs = 0
        sim.step(1)
        o = sim.store()
        np.testing.assert_array_equal(g["ids"], o["ids"])
        np.testing.assert_array_equal(g["tags"], o["tags"])
        np.testing.assert_array_equal(g["finite_volumes"], o["finite_volumes"])
        np.testing.assert_array_equal(g["boundary_measures"], o["boundary_measures"])
        assert tuple(gpu.current_mcs_report()) == sim.report()
        words = np.where(g["tags"] == 1, g["ids"].astype(np.int64), -g["ids"].astype(np.int64))
        ident, scale = 0, 1
        for w in words:
            ident += labels.index(int(w)) * scale
            scale *= len(labels)
        counts[ident] += 1
    assert counts.sum() == 512


@pytest.mark.parametrize("alg_type", [CheckerboardSweepCPM, LotteryCPM])
@pytest.mark.parametrize("dims", [(4, 4), (4, 4, 4), (3, 3), (5, 7)])
def test_reference_fixture(gpu_lib, alg_type, dims):
    compiled, proposal, components = scientific_fixture(dims)
    gpu, sim, reps = run_parity(compiled, proposal, components, alg_type(np.float32(5.0)), 0xC01A,
                                [1, 2, 5, 20])
    for rep in reps:
        _partition(rep)
    if alg_type is CheckerboardSweepCPM:
        assert reps[0].activated_attempts == int(np.prod(dims))
        if dims == (3, 3):
            assert reps[0].internal_rounds >= 3


@pytest.mark.parametrize("alg", ["checkerboard", "lottery"])
def test_config1_plain(gpu_lib, alg):
    compiled, proposal, components, algorithm, seed = workloads.config1_plain(algorithm=alg)
    gpu, sim, reps = run_parity(compiled, proposal, components, algorithm, seed, [1, 10, 25, 100])
    for rep in reps:
        _partition(rep)
    # periodic and obstacle-free: the lottery kernel that decides winners from tickets alone
    assert bool(gpu.kernel_config() & 64) == (alg == "lottery")


@pytest.mark.parametrize("alg", ["checkerboard", "lottery"])
def test_scaled_with_boundary_energy(gpu_lib, alg):
    compiled, proposal, components, algorithm, seed = workloads.scaled_2d(300, algorithm=alg, seed=99)
    run_parity(compiled, proposal, components, algorithm, seed, [1, 8])


@pytest.mark.parametrize("alg", ["checkerboard", "lottery"])
def test_cubes_3d_float_contact(gpu_lib, alg):
    compiled, proposal, components, algorithm, seed = workloads.cubes_3d(L=24, side=6, algorithm=alg,
                                                                         seed=7)
    run_parity(compiled, proposal, components, algorithm, seed, [1, 4])


@pytest.mark.parametrize("dim", [2, 3])
def test_lottery_moore_shell_winner_test(gpu_lib, dim):
    """Plain lotteries whose conflict shell is the radius-1 Moore shell (Moore or 26-neighbour
    contact over a von Neumann proposal) take the winner test over compile-time ticket offsets
    (kernel_config bit 7); results equal the oracle."""
    if dim == 2:
        model = workloads.scaled_2d(130, algorithm="lottery", seed=5)
    else:
        model = workloads.cubes_3d(L=24, side=6, algorithm="lottery", seed=5)
    gpu, sim, reps = run_parity(*model, [1, 4])
    assert gpu.kernel_config() & 64 and gpu.kernel_config() & 128
    for rep in reps:
        _partition(rep)


def _walled_domain(alg_type, obstacles=True):
    L = 24
    wall = MediumOwner(7)
    obstacle_list = [((5, 5), MediumOwner(9)), ((6, 5), MediumOwner(9)), ((12, 20), MediumOwner(9))] if obstacles else []
    domain = CartesianDomain((L, L), boundaries=(AxisBoundary(FixedExterior(wall), ClosedBoundary()),
                                                 AxisBoundary(ClosedBoundary())),
                             obstacles=obstacle_list)
    owners = np.full((L, L), MediumOwner(1), dtype=np.int64)
    cid = 0
    types = {}
    for cx in range(0, L - 4, 6):
        for cy in range(0, L - 4, 6):
            cid += 1
            owners[cx + 1:cx + 5, cy + 1:cy + 5] = cid
            types[cid] = 2 + cid % 2
    for (x, y), o in obstacle_list:
        owners[x - 1, y - 1] = o
    proposal = first_shell_relation(ProposalRole, 2)
    surface = first_shell_relation(SurfaceRole, 2)
    contact_rel = moore_relation(ContactRole, 2)
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    J = np.array([[0, 9, 9], [9, 3, 7], [9, 7, 4]], dtype=np.float32)
    contact = UnorderedContactHamiltonian(J, MediumTypeTable({1: 1, 7: 1, 9: 1}), contact_rel)
    boundary = QuadraticBoundaryHamiltonian(BoundaryEdgeCount(), surface, number_type=np.float32)
    schema = {}
    for c in (volume, boundary):
        schema.update(required_properties(c))
    state = LogicalPottsState(owners, cid + 2, cell_types=types, medium_domains=(1, 7, 9),
                              property_schema=schema)
    state.property_values("target_volume")[:cid] = 16
    state.property_values("volume_strength")[:cid] = 2
    state.property_values("target_boundary")[:cid] = 16
    state.property_values("boundary_strength")[:cid] = 1
    compiled = compile_scientific_state(state, domain, BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    return compiled, proposal, ScientificComponentSet(energies=(volume, contact, boundary)), alg_type(np.float32(6.0))


@pytest.mark.parametrize("alg_type", [CheckerboardSweepCPM, LotteryCPM])
@pytest.mark.parametrize("obstacles", [False, True])
def test_closed_fixed_obstacle_domain(gpu_lib, alg_type, obstacles):
    compiled, proposal, components, alg = _walled_domain(alg_type, obstacles)
    gpu, _, _ = run_parity(compiled, proposal, components, alg, 31337, [1, 5, 15])
    assert not gpu.kernel_config() & 64  # walls: the lottery kernel that checks every neighbour


def test_replay_and_metrics(gpu_lib):
    compiled, proposal, components, algorithm, seed = workloads.config1_plain()
    a = init_scientific(compiled, proposal, components, algorithm, seed=seed)
    b = init_scientific(compiled, proposal, components, algorithm, seed=seed)
    assert a.current_mcs_report() is None
    a.step(3)
    b.step(3)
    before = a.metrics()
    ra = a.current_mcs_report()
    after = a.metrics()
    assert after["host_synchronizations"] > before["host_synchronizations"]
    assert ra == b.current_mcs_report()
    np.testing.assert_array_equal(a.download()["ids"], b.download()["ids"])
    count, order = a.schedule()
    assert sorted(order.tolist()) == list(range(1, count + 1))


def test_graph_chunks_match_plain_launches(gpu_lib):
    compiled, proposal, components, algorithm, seed = workloads.config1_plain()
    a = init_scientific(compiled, proposal, components, algorithm, seed=seed)
    b = init_scientific(compiled, proposal, components, algorithm, seed=seed)
    b.set_graph_chunk(5)
    a.step(20)
    b.step(20)
    for _ in range(3):  # replay with host work between replays
        a.step(5)
        b.step(5)
        assert a.current_mcs_report() == b.current_mcs_report()
    np.testing.assert_array_equal(a.download()["ids"], b.download()["ids"])
    assert b.metrics()["graph_launches"] >= 4


def test_prepared_graph_matches_plain_launches(gpu_lib):
    """A graph captured ahead of the steps (cpm_prepare_graph performs no MCS) replays the same
    MCS as plain launches, after a warmup whose length is not a multiple of the chunk."""
    compiled, proposal, components, algorithm, seed = workloads.config1_plain()
    a = init_scientific(compiled, proposal, components, algorithm, seed=seed)
    b = init_scientific(compiled, proposal, components, algorithm, seed=seed)
    a.step(5)
    b.step(5)
    b.set_graph_chunk(10)
    assert b.prepare_graph()
    assert b.metrics()["graph_launches"] == 0
    assert b.current_mcs_report() == a.current_mcs_report()  # nothing performed
    a.step(20)
    b.step(20)
    assert b.metrics()["graph_launches"] == 2
    assert a.current_mcs_report() == b.current_mcs_report()
    for f in ("ids", "tags", "finite_volumes", "boundary_measures"):
        np.testing.assert_array_equal(a.download()[f], b.download()[f], err_msg=f)
    b.set_graph_chunk(0)
    assert not b.prepare_graph()
