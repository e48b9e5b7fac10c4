"""Every compiled tile shape gives the oracle's results bit for bit (the shape only changes
launch geometry): 2D checkerboard and lottery on odd and even sides, and 3D float-J cubes."""
import numpy as np
import pytest

from corepotts import init_scientific, set_tile_shape, tile_shapes, workloads
from corepotts.spec import ALG_CHECKERBOARD, ALG_LOTTERY
from helpers import oracle_for

pytestmark = pytest.mark.gpu

FIELDS = ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures", "active",
          "cell_types")


def _parity(model, mcs):
    compiled, proposal, components, alg, seed = model
    gpu = init_scientific(compiled, proposal, components, alg, seed=seed)
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
