"""Oracle pins beyond the transition goldens:

* the reference's literal cell-claim fixtures (lib/CorePotts/test/test_checkerboard_algorithms.jl:180-196);
* the single 64-bit key the HIP kernels use for claims, (priority << 32) | ~recipient resolved by
  atomicMax, selects exactly what the reference's three-phase protocol selects
  (max priority, then min recipient id among the max holders, checkerboard.jl:273-369);
* canonical direction order (lib/CorePotts/test/test_cartesian_relations.jl:25-32);
* LotteryCPM's activation statistics, restated from lib/CorePotts/test/test_lottery_algorithms.jl:78-181
  (the reference holds no golden vectors for the lottery sweep).
"""
import numpy as np
import pytest

from corepotts import (AxisBoundary, BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain,
                       ClosedBoundary, ContactRole, LogicalPottsState, LotteryCPM, MediumOwner,
                       ProposalRole, ScientificComponentSet, SurfaceRole, build_spec,
                       compile_scientific_state, first_shell_relation, moore_relation,
                       state_arrays)
from oracle import bridge


def test_reference_claim_fixtures():
    # chain: attempt k moves recipient k from cell k to cell k+1
    sel, disp, maxp, mini = bridge.claims([1, 2, 3], [1, 2, 3], [2, 3, 4], [10, 20, 30], 8)
    assert list(sel) == [0, 0, 1]
    assert list(disp) == [1, 1, 0]
    assert list(maxp[2:4]) == [20, 30]
    # exact priority tie on cell 2: the smaller recipient id (3) wins
    sel, disp, maxp, mini = bridge.claims([5, 3], [1, 2], [2, 3], [7, 7], 8)
    assert list(sel) == [0, 1]
    assert mini[2] == 3


def _key_select(recipient, losing, gaining, priority):
    """The kernels' formulation: one 64-bit maximum per cell, media (id 0) never claimed."""
    key = (priority.astype(np.uint64) << np.uint64(32)) | (~recipient.astype(np.uint32)).astype(np.uint64)
    best = {}
    for cells in (losing, gaining):
        for c, k in zip(cells, key):
            if c:
                best[c] = max(best.get(c, 0), int(k))
    return np.array([(l == 0 or best[l] == int(k)) and (g == 0 or best[g] == int(k))
                     for l, g, k in zip(losing, gaining, key)], dtype=np.uint8)


@pytest.mark.parametrize("seed", range(40))
def test_atomic_key_matches_three_phase_claims(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 60))
    cap = int(rng.integers(2, 10))
    recipient = rng.permutation(np.arange(1, 4 * n + 1))[:n].astype(np.uint32)
    losing = rng.integers(0, cap + 1, size=n).astype(np.uint32)
    gaining = rng.integers(0, cap + 1, size=n).astype(np.uint32)
    same = losing == gaining
    gaining[same] = (gaining[same] % cap) + 1  # actionable: owners differ
    hi = int(rng.choice([3, 50, 2 ** 32 - 1]))  # small ranges force priority ties
    priority = rng.integers(0, hi, size=n, endpoint=True).astype(np.uint32)
    if seed % 5 == 0:
        priority[:] = np.uint32(0xFFFFFFFF)
    sel, disp, _, _ = bridge.claims(recipient, losing, gaining, priority, cap)
    np.testing.assert_array_equal(_key_select(recipient, losing, gaining, priority), sel)
    assert np.all((disp == 1) == (sel == 0))


def test_canonical_direction_order():
    assert first_shell_relation(ProposalRole, 2).offsets == ((-1, 0), (0, -1), (0, 1), (1, 0))
    assert moore_relation(ContactRole, 2).offsets[4:] == ((-1, -1), (-1, 1), (1, -1), (1, 1))
    # physical spacing makes -y the first canonical direction
    assert first_shell_relation(ProposalRole, 2, spacing=(2.0, 1.0)).offsets == \
        ((0, -1), (0, 1), (-1, 0), (1, 0))
    three = first_shell_relation(ProposalRole, 3).offsets
    assert three == ((-1, 0, 0), (0, -1, 0), (0, 0, -1), (0, 0, 1), (0, 1, 0), (1, 0, 0))
    assert len(moore_relation(ContactRole, 3).offsets) == 26


def _medium_lattice(dims, boundaries=None, seed=0):
    state = LogicalPottsState(np.full(int(np.prod(dims)), MediumOwner(1), dtype=np.int64), 1,
                              medium_domains=(1,), dims=dims)
    domain = CartesianDomain(dims) if boundaries is None else CartesianDomain(dims, boundaries=boundaries)
    proposal = first_shell_relation(ProposalRole, len(dims))
    tracker = BoundaryMeasureTracker(BoundaryEdgeCount(), first_shell_relation(SurfaceRole, len(dims)))
    compiled = compile_scientific_state(state, domain, tracker)
    spec = build_spec(compiled, proposal, ScientificComponentSet(), LotteryCPM(np.float32(1.0)), seed)
    return bridge.OracleSim(spec, state_arrays(compiled, spec)), dims


def test_lottery_calibration_uniform_across_closed_boundaries():
    # test_lottery_algorithms.jl:78-114
    sim, dims = _medium_lattice((5, 5), (AxisBoundary(ClosedBoundary()), AxisBoundary(ClosedBoundary())),
                                seed=0xCA11)
    assert sim.rounds() == 5
    mcs_count = 200
    counts = np.zeros(25, np.int64)
    orders = set()
    for _ in range(mcs_count):
        sim.step(1)
        order = sim.last_order()
        assert sorted(order) == [1, 2, 3, 4, 5]
        orders.add(tuple(order))
        counts += sim.last_activations()
    assert len(orders) > 1
    assert np.all(np.abs(counts - mcs_count) <= 50)
    corners = counts[[0, 4, 20, 24]]
    interior = counts[[6, 7, 8, 11, 12, 13, 16, 17, 18]]
    assert abs(corners.mean() - interior.mean()) <= 30


def test_lottery_waiting_repetition_and_spatial_statistics():
    # test_lottery_algorithms.jl:130-181
    sim, dims = _medium_lattice((8, 8), seed=0x57A71571C)
    assert sim.rounds() == 5
    mcs_count, sites = 600, 64
    activated = np.zeros((mcs_count, sites), np.int64)
    for m in range(mcs_count):
        sim.step(1)
        activated[m] = sim.last_activations()
    samples = activated.ravel()
    expected_zero = (4 / 5) ** 5
    expected_repeat = 1 - expected_zero - (4 / 5) ** 4
    assert abs(samples.mean() - 1) <= 0.03
    assert abs(samples.var() - 4 / 5) <= 0.04
    assert abs((samples == 0).mean() - expected_zero) <= 0.02
    assert abs((samples >= 2).mean() - expected_repeat) <= 0.02
    waits = []
    for s in range(sites):
        active = np.flatnonzero(activated[:, s] > 0)
        waits.extend(np.diff(active))
    assert abs(np.mean(waits) - 1 / (1 - expected_zero)) <= 0.06
    grid = np.arange(sites).reshape(dims, order="F")
    products = []
    for y in range(dims[1]):
        for x in range(dims[0]):
            left, right = grid[x, y], grid[(x + 1) % dims[0], y]
            products.append((activated[:, left] - 1) * (activated[:, right] - 1))
    # adjacent sites are never local ticket maxima in the same round: covariance -1/5 per MCS
    assert abs(np.mean(np.concatenate(products)) + 1 / 5) <= 0.04
