"""The oracle against the reference's realistic-scale trajectories.

design/evidence/phase-13/realistic/cpu/single_cell_migration-{SequentialCPM,
CheckerboardSweepCPM}.toml (extracted by tests/golden/extract_realistic_migration.py) hold,
for 512 replicas each, what the reference's Julia 1.12.6 CPU run of 2,500 MCS produced: the
terminal and mean energy per site, the mean boundary measure and the mixing time (exact
values: the energies are integers over 4096 sites), and the centre-of-mass displacement
statistics (floating point through sin, cos and atan). A replica's dynamics seed fixes its
whole trajectory (Philox addresses, colouring, colour order, claims, integer energy changes,
the Float32 acceptance exponential and uniform draw), so equal observables after 2,500 MCS
pin the oracle over long runs, including every acceptance decision of the trajectory.

By default a spread of CPM_MIGRATION_REPLICAS (24) replicas per algorithm runs here; set it to
512 to check every replica (about 2 minutes on 8 cores).
"""
import multiprocessing
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np
import pytest

from helpers import (EXACT_MIGRATION_FIELDS, FLOAT_MIGRATION_FIELDS, load_migration,
                     oracle_migration_replica)

ALGORITHMS = {"sequential": "SequentialCPM", "checkerboard": "CheckerboardSweepCPM"}


def _replicas(total=512):
    n = min(total, int(os.environ.get("CPM_MIGRATION_REPLICAS", "24")))
    return sorted({int(round(i * (total - 1) / max(n - 1, 1))) for i in range(n)})


def _run(args):
    algorithm, seed = args
    return oracle_migration_replica(algorithm, seed)


def test_replica_seed_derivation():
    """derive_replica_seeds (RealisticScaleRunner.jl:44-59): dynamics seed = splitmix64 of the
    replica seed xor 0x13d90113d90113d9; replica seeds are 0x1301000000000000 + index."""
    def splitmix64(v):
        m = (1 << 64) - 1
        v = (v + 0x9e3779b97f4a7c15) & m
        v = ((v ^ (v >> 30)) * 0xbf58476d1ce4e5b9) & m
        v = ((v ^ (v >> 27)) * 0x94d049bb133111eb) & m
        return v ^ (v >> 31)
    data = load_migration()
    for name in ALGORITHMS.values():
        col = data["algorithms"][name]
        for i, (seed, dyn) in enumerate(zip(col["seed_hex"], col["dynamics_seed_hex"])):
            assert int(seed, 16) == 0x1301000000000000 + i
            assert splitmix64(int(seed, 16) ^ 0x13d90113d90113d9) == int(dyn, 16)


@pytest.mark.parametrize("algorithm", list(ALGORITHMS))
def test_oracle_reproduces_realistic_trajectories(algorithm):
    col = load_migration()["algorithms"][ALGORITHMS[algorithm]]
    idx = _replicas()
    jobs = [(algorithm, int(col["dynamics_seed_hex"][i], 16)) for i in idx]
    # spawned workers: a forked child of a process whose OpenMP thread pool already exists
    # (earlier oracle tests in the same process) can crash in libgomp
    ctx = multiprocessing.get_context("spawn")
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1), mp_context=ctx) as pool:
        results = list(pool.map(_run, jobs))
    mismatches = []
    for i, obs in zip(idx, results):
        for key in EXACT_MIGRATION_FIELDS:
            if obs[key] != col[key][i]:
                mismatches.append((i, key, obs[key], col[key][i]))
        for key in FLOAT_MIGRATION_FIELDS:
            if not np.isclose(obs[key], col[key][i], rtol=1e-9, atol=1e-12):
                mismatches.append((i, key, obs[key], col[key][i]))
    assert not mismatches, f"{len(mismatches)} mismatches, first {mismatches[:5]}"
