"""FP64 energy bookkeeping on the GPU (cpm_report_energy, cpm_total_energy) against the oracle.

* cpm_total_energy (global FP64 energy of the device state) equals the oracle's
  orc_total_energy on the same state: exactly for integer-valued models, to 1e-12 relative for
  the Float32 3D cubes (the two sum in different orders).
* The per-MCS FP64 sums of the accepted copies' delta_h (cpm_report_energy) equal the oracle's
  sums MCS by MCS, and add up to the change of the global energy: exactly for integer models,
  within the Float32 fold rounding bound for the cubes (see tests/test_oracle_drift.py).
"""
import numpy as np
import pytest

from corepotts import init_scientific, workloads
from helpers import oracle_for


def _pair(model):
    gpu = init_scientific(*model)
    sim, _ = oracle_for(*model)
    return gpu, sim


@pytest.mark.gpu
@pytest.mark.parametrize("name,model_fn,mcs", [
    ("config1 checkerboard", lambda: workloads.config1_plain(), 10),
    ("config1 lottery", lambda: workloads.config1_plain(algorithm="lottery"), 5),
    ("config2 4096^2 checkerboard", lambda: workloads.config2_large(), 4),
    ("config2 rule 1024^2 lottery", lambda: workloads.scaled_2d(1024, algorithm="lottery"), 2),
])
def test_integer_models_energy_exact(name, model_fn, mcs):
    gpu, sim = _pair(model_fn())
    e0 = gpu.total_energy()
    assert e0 == sim.total_energy()
    total = 0.0
    for _ in range(mcs):
        gpu.step(1)
        sim.step(1)
        dh = gpu.accepted_energy()
        assert dh == sim.accepted_energy()[0]
        total += dh
    e1 = gpu.total_energy()
    assert e1 == sim.total_energy()
    assert total == e1 - e0
    assert total != 0.0
    gpu.close()


@pytest.mark.gpu
def test_cubes_float32_energy_against_oracle():
    gpu, sim = _pair(workloads.cubes_3d(128))
    e0 = gpu.total_energy()
    assert e0 == pytest.approx(sim.total_energy(), rel=1e-12)
    total = 0.0
    for _ in range(5):
        gpu.step(1)
        sim.step(1)
        dh = gpu.accepted_energy()
        assert dh == pytest.approx(sim.accepted_energy()[0], rel=1e-12, abs=1e-9)
        total += dh
    assert gpu.total_energy() == pytest.approx(sim.total_energy(), rel=1e-12)
    gpu.close()


@pytest.mark.gpu
def test_cubes_256_drift_is_rounding_only():
    """256^3 (the 3D configuration), 20 MCS: the FP64 sum of accepted delta_h tracks the
    global energy change within the Float32 fold rounding bound."""
    gpu = init_scientific(*workloads.cubes_3d(256))
    e0 = gpu.total_energy()
    total, accepted = 0.0, 0
    for _ in range(20):
        gpu.step(1)
        total += gpu.accepted_energy()
        accepted += gpu.current_mcs_report().accepted_copies
    change = gpu.total_energy() - e0
    eps32 = float(np.finfo(np.float32).eps)
    largest_fold = 2 * 26 * 20.0 + 2 * 1.5 * 64.0 ** 2 + 2 * 0.5 * 512.0 ** 2
    assert accepted > 100000
    assert abs(total - change) <= accepted * 60 * eps32 / 2 * largest_fold
    assert abs(total - change) < accepted * 0.01
    gpu.close()
