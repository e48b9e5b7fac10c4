"""Energy bookkeeping of the oracle at scale: the FP64 sum of the accepted copies' delta_h
equals the change of the global energy (orc_total_energy, FP64).

Each accepted copy changes the global energy by exactly its energy change (the reference's
local/global identity, lib/CorePotts/test/test_scientific_hamiltonians.jl:47-80), and copies
of one colour pass or lottery round touch disjoint cells and non-adjacent sites, so their
changes add. For integer-valued models (the 200^2 and 4096^2 workloads: integer J, lambda
and targets) every delta_h is an exact integer and the identity holds exactly. For the Float32
3D cube workload (random J in [2, 20], lambda_V = 1.5, lambda_S = 0.5) each delta_h carries
the rounding of its Float32 fold; the drift is bounded by the accepted count times
60 eps32 / 2 times the largest sum of absolute terms of a fold (26 contact directions at
J <= 20 on both sides, plus the volume and boundary terms), while a wrong, dropped or
doubled term would move the sum by about a J per copy.
"""
import numpy as np
import pytest

from corepotts import build_spec, state_arrays, workloads
from oracle.bridge import OracleSim


def _sim(model):
    spec = build_spec(*model)
    return OracleSim(spec, state_arrays(model[0], spec))


@pytest.mark.parametrize("algorithm,mcs", [("checkerboard", 30), ("lottery", 10)])
def test_config1_accepted_energy_equals_global_change(algorithm, mcs):
    sim = _sim(workloads.config1_plain(algorithm=algorithm))
    e0 = sim.total_energy()
    accepted = 0
    last_sum = 0.0
    for _ in range(mcs):
        sim.step(1)
        last, _total = sim.accepted_energy()
        last_sum += last
        accepted += sim.report()[11]
    _, total = sim.accepted_energy()
    assert total == last_sum
    assert accepted > 0
    assert total == sim.total_energy() - e0
    assert total < 0  # the seeded squares relax


def test_config2_full_size_accepted_energy_equals_global_change():
    """4096^2, 163,840 cells, volume + Moore contact + boundary (integer), 5 MCS."""
    sim = _sim(workloads.config2_large())
    e0 = sim.total_energy()
    sim.step(5)
    _, total = sim.accepted_energy()
    assert total == sim.total_energy() - e0


def test_config2_rule_sequential_accepted_energy_equals_global_change():
    model = workloads.scaled_2d(256)
    spec = build_spec(*model)
    seq = OracleSim(spec, state_arrays(model[0], spec), 0)  # SequentialCPM
    e0 = seq.total_energy()
    seq.step(5)
    assert seq.accepted_energy()[1] == seq.total_energy() - e0


@pytest.mark.parametrize("L,mcs", [(64, 10), (128, 5)])
def test_config3_float32_drift_is_rounding_only(L, mcs):
    sim = _sim(workloads.cubes_3d(L))
    e0 = sim.total_energy()
    accepted = 0
    for _ in range(mcs):
        sim.step(1)
        accepted += sim.report()[11]
    _, total = sim.accepted_energy()
    change = sim.total_energy() - e0
    eps32 = float(np.finfo(np.float32).eps)
    largest_fold = 2 * 26 * 20.0 + 2 * 1.5 * 64.0 ** 2 + 2 * 0.5 * 512.0 ** 2
    bound = accepted * 60 * eps32 / 2 * largest_fold
    assert accepted > 1000
    assert abs(total - change) <= bound, (total, change, bound)
    # the sum tracks the change to far better than one J per hundred copies
    assert abs(total - change) < accepted * 0.01
