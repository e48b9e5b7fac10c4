"""The engine's realized greedy colouring (cpm_model_colors, computed on the host as cpm_create
computes it) against the oracle's site-by-site greedy colouring (checkerboard.jl:57-90),
which the golden fixtures pin. Obstacle-free lattices take the layer recurrence (a layer is a
function of the layers before it, so a repeated state repeats the layers that follow); these
cases cover periods that do and do not divide the lattice, seams, closed faces, radius-2
relations, lattices too short for a repeat, and 3D planes."""
import numpy as np
import pytest

from corepotts import (CheckerboardSweepCPM, ContactRole, MediumTypeTable, ScientificComponentSet,
                       UnorderedContactHamiltonian, build_spec, static_relation)
from corepotts.native import model_colors
from oracle import bridge
from test_oracle_tiled import _tissue


def _check(compiled, proposal, components):
    spec = build_spec(compiled, proposal, components, CheckerboardSweepCPM(np.float32(5.0)), 1)
    colors, count = model_colors(spec)
    expected = bridge.coloring(spec)
    np.testing.assert_array_equal(colors, expected)
    assert count == int(expected.max())
    return count


@pytest.mark.parametrize("dims", [(24, 20), (23, 17), (5, 5), (6, 7), (31, 31), (64, 33), (33, 64), (40, 3),
                                  (3, 40), (97, 45)])
@pytest.mark.parametrize("contact", ["moore", "vn"])
@pytest.mark.parametrize("periodic", [True, False], ids=["periodic", "closed"])
def test_2d_colourings_equal_the_oracle(dims, contact, periodic):
    _check(*_tissue(dims, 3, None, contact, True, periodic))


@pytest.mark.parametrize("dims", [(9, 10, 11), (12, 12, 12), (7, 7, 30), (5, 6, 5), (16, 9, 13)])
@pytest.mark.parametrize("periodic", [True, False], ids=["periodic", "closed"])
def test_3d_colourings_equal_the_oracle(dims, periodic):
    _check(*_tissue(dims, 4, None, "moore", True, periodic))


@pytest.mark.parametrize("dims", [(13, 8), (30, 29), (41, 40), (9, 9), (25, 50)])
@pytest.mark.parametrize("periodic", [True, False], ids=["periodic", "closed"])
def test_radius_two_colourings_equal_the_oracle(dims, periodic):
    compiled, proposal, components = _tissue(dims, 11, None, "vn", False, periodic)
    rel = static_relation(ContactRole, [(1, 0), (-1, 0), (0, 1), (0, -1), (2, 0), (-2, 0), (0, 2), (0, -2),
                                        (1, 1), (-1, -1)])
    contact = UnorderedContactHamiltonian(np.array([[0, 5, 7], [5, 2, 6], [7, 6, 3]], np.float32),
                                          MediumTypeTable({1: 1}), rel)
    _check(compiled, proposal, ScientificComponentSet(energies=(components.energies[0], contact)))


def test_walls_and_obstacles_colouring_equals_the_oracle():
    from test_gpu_parity import _walled_domain
    compiled, proposal, components, _ = _walled_domain(CheckerboardSweepCPM, obstacles=True)
    _check(compiled, proposal, components)


def test_large_periodic_lattice_takes_a_few_colours():
    # 1000 x 998 periodic, Moore conflicts: the interior repeats a 2 x 2 pattern and the
    # seams follow from it; the recurrence must reproduce the oracle everywhere
    assert _check(*_tissue((1000, 998), 5, None, "moore", True, True)) == 4
    assert _check(*_tissue((999, 1001), 5, None, "moore", True, True)) >= 4


def _obstacle_domain(dims, count, seed, periodic=True):
    """Medium-filled lattice with `count` obstacle sites (medium 9) at random positions."""
    from corepotts import (AxisBoundary, BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain,
                           ClosedBoundary, LogicalPottsState, MediumOwner, PeriodicBoundary, ProposalRole,
                           QuadraticVolumeHamiltonian, SurfaceRole, compile_scientific_state,
                           first_shell_relation, moore_relation, required_properties)
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    picks = rng.choice(n, size=count, replace=False)
    coords = [tuple(int(c) + 1 for c in np.unravel_index(p, dims, order="F")) for p in picks]
    b = PeriodicBoundary() if periodic else ClosedBoundary()
    domain = CartesianDomain(dims, boundaries=tuple(AxisBoundary(b) for _ in dims),
                             obstacles=[(c, MediumOwner(9)) for c in coords])
    owners = np.full(dims, MediumOwner(1), dtype=np.int64)
    owners[tuple(slice(0, 2) for _ in dims)] = 1
    for c in coords:
        owners[tuple(k - 1 for k in c)] = MediumOwner(9)
    ndim = len(dims)
    volume = QuadraticVolumeHamiltonian(number_type=np.float32)
    contact = UnorderedContactHamiltonian(np.array([[0, 5], [5, 2]], np.float32), MediumTypeTable({1: 1, 9: 1}),
                                          moore_relation(ContactRole, ndim))
    state = LogicalPottsState(owners, 2, cell_types={1: 2}, medium_domains=(1, 9),
                              property_schema=required_properties(volume))
    surface = first_shell_relation(SurfaceRole, ndim)
    compiled = compile_scientific_state(state, domain, BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    return compiled, first_shell_relation(ProposalRole, ndim), ScientificComponentSet(energies=(volume, contact))


@pytest.mark.parametrize("dims,count", [((200, 150), 5), ((120, 301), 40), ((64, 64), 1), ((33, 35), 200),
                                        ((20, 18, 40), 12)])
@pytest.mark.parametrize("periodic", [True, False], ids=["periodic", "closed"])
def test_obstacle_colourings_equal_the_oracle(dims, count, periodic):
    _check(*_obstacle_domain(dims, count, 9 + count, periodic))
