"""Parity at the full sizes of the benchmark configurations: the HIP engine and the oracle
start from the same 4096 x 4096 (163,840 cells) and 256^3 (32,768 cubes, Float32 contact
matrix) states and must agree bit for bit on lattice, trackers and report after the
compared MCS (checkerboard, lottery and tiled at 4096^2); the device state must also pass
the tracker recount."""
import numpy as np
import pytest

from corepotts import init_scientific, tracker_conformance_errors, workloads
from helpers import oracle_for

pytestmark = pytest.mark.gpu

FIELDS = ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures", "active",
          "cell_types")


def _full(model, mcs):
    compiled, proposal, components, alg, seed = model
    gpu = init_scientific(compiled, proposal, components, alg, seed=seed)
    sim, _ = oracle_for(compiled, proposal, components, alg, seed)
    gpu.step(mcs)
    sim.step(mcs)
    g, o = gpu.download(), sim.store()
    for f in FIELDS:
        np.testing.assert_array_equal(g[f], o[f], err_msg=f)
    assert tuple(gpu.current_mcs_report()) == sim.report()
    assert not tracker_conformance_errors(gpu.logical_state())
    gpu.close()


def test_config2_checkerboard_4096(gpu_lib):
    _full(workloads.config2_large(), 2)


def test_config2_lottery_4096(gpu_lib):
    _full(workloads.config2_large(algorithm="lottery"), 1)


def test_config3_cubes_256(gpu_lib):
    _full(workloads.cubes_3d(256), 1)


def test_config2_tiled_4096(gpu_lib):
    # TiledCheckerboardCPM with the reference's default 4 x 4 tiles on the config-2 state
    _full(workloads.config2_large(algorithm="tiled"), 2)
