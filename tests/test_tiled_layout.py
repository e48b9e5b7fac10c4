"""The engine's TiledCheckerboardCPM layout (cpm_tiled_layout, computed on the host) against
the oracle's, which the oracle tests pin to the reference's all-pairs greedy colouring
(tiled_checkerboard.jl:131-188), and the layouts the engine refuses: tiles of one colour
that a partial tile at a periodic seam brings within the halo of each other."""
import numpy as np
import pytest

from corepotts import (ContactRole, MediumTypeTable, ScientificComponentSet, TiledCheckerboardCPM,
                       UnorderedContactHamiltonian, build_spec, state_arrays, static_relation)
from corepotts.native import EngineError, tiled_layout
from oracle import bridge
from test_oracle_tiled import _tissue, reference_fixture

CPM_E_UNSUPPORTED = 2


def _both(compiled, proposal, components, algorithm, seed=7):
    spec = build_spec(compiled, proposal, components, algorithm, seed)
    sim = bridge.OracleSim(spec, state_arrays(compiled, spec))
    return tiled_layout(spec), sim


@pytest.mark.parametrize("case", [
    ((24, 20), 1, "moore", True, True, None, None),
    ((23, 17), 2, "vn", True, False, (4, 3), 5),
    ((30, 30), 3, "moore", False, True, (3, 5), None),
    ((10, 9, 8), 4, "moore", True, True, None, 3),
    ((9, 8, 7), 5, "vn", True, False, (2, 3, 2), None),
    ((11, 11), 6, "vn", True, True, (3, 3), 4),
    ((16, 12), 7, "moore", True, True, (1, 1), None),
], ids=["2d-default", "2d-closed-4x3", "2d-3x5", "3d-default", "3d-closed", "2d-seam-3x3", "2d-1x1"])
def test_engine_layout_matches_the_oracle(case):
    dims, seed, contact, boundary, periodic, tile, interval = case
    compiled, proposal, components = _tissue(dims, seed, None, contact, boundary, periodic)
    algorithm = TiledCheckerboardCPM(np.float32(12.0), tile_size=tile, switching_interval=interval)
    eng, sim = _both(compiled, proposal, components, algorithm)
    grid, reach, halo, colors, count = sim.tiled_layout()
    assert eng["grid"] == tuple(grid) and eng["reach"] == tuple(reach) and eng["halo"] == halo
    np.testing.assert_array_equal(eng["colors"], colors)
    assert eng["color_count"] == count
    # internal rounds of one MCS: colours x batches (tiled_checkerboard_device.jl:790-806)
    sim.step(1)
    assert sim.report()[1] == eng["color_count"] * eng["batches"]
    edges = tile or ((4, 4) if len(dims) == 2 else (2, 2, 2))
    # the staged kernel holds a box plus halo of at most 64 positions
    box = int(np.prod([e + 2 * halo for e in edges]))
    assert eng["staged"] == (box if box <= 64 else 0)
    tile_sites = int(np.prod(edges))  # every tile can be full
    expected_interval = min(interval, tile_sites) if interval else tile_sites
    assert eng["interval"] == expected_interval


def test_reference_fixture_layout():
    # test_tiled_checkerboard.jl:39-90: 6 x 6, tile (2, 2), interval 2 -> two batches per colour
    compiled, proposal, components = reference_fixture()
    eng, sim = _both(compiled, proposal, components,
                     TiledCheckerboardCPM(np.float32(5.0), tile_size=(2, 2), switching_interval=2))
    assert eng["grid"][:2] == (3, 3) and eng["halo"] == 1 and eng["interval"] == 2 and eng["batches"] == 2
    np.testing.assert_array_equal(eng["colors"], sim.tiled_layout()[3])


def _radius_two(dims, tile, periodic=True):
    """Contact relation reaching distance 2 on a tissue; halo radius 2."""
    compiled, proposal, components = _tissue(dims, 11, None, "vn", False, periodic)
    rel = static_relation(ContactRole, [(1, 0), (-1, 0), (0, 1), (0, -1), (2, 0), (-2, 0), (0, 2), (0, -2)])
    contact = UnorderedContactHamiltonian(np.array([[0, 5, 7], [5, 2, 6], [7, 6, 3]], np.float32),
                                          MediumTypeTable({1: 1}), rel)
    components = ScientificComponentSet(energies=(components.energies[0], contact))
    return compiled, proposal, components, TiledCheckerboardCPM(np.float32(8.0), tile_size=tile)


def test_periodic_seam_within_the_halo_is_refused():
    # x: 13 sites in tiles of 4 -> tiles 0 and 2 share a colour (tile separation 2 > reach 1),
    # but across the seam their sites are 2 apart, within the halo radius 2
    compiled, proposal, components, algorithm = _radius_two((13, 8), (4, 4))
    spec = build_spec(compiled, proposal, components, algorithm, 3)
    sim = bridge.OracleSim(spec, state_arrays(compiled, spec))
    grid, reach, halo, colors, _ = sim.tiled_layout()
    assert halo == 2 and reach[0] == 1 and colors[0] == colors[2]
    with pytest.raises(EngineError) as err:
        tiled_layout(spec)
    assert err.value.code == CPM_E_UNSUPPORTED and "seam" in str(err.value)


@pytest.mark.parametrize("dims,tile,periodic", [((16, 8), (4, 4), True), ((13, 8), (4, 4), False),
                                                ((14, 9), (4, 4), True)])
def test_seam_check_accepts_safe_layouts(dims, tile, periodic):
    compiled, proposal, components, algorithm = _radius_two(dims, tile, periodic)
    eng, sim = _both(compiled, proposal, components, algorithm)
    np.testing.assert_array_equal(eng["colors"], sim.tiled_layout()[3])
    assert eng["halo"] == 2
