"""HIP CheckerboardSweepCPM against the reference's realistic-scale trajectories.

All 512 single-cell-migration replicas of
design/evidence/phase-13/realistic/cpu/single_cell_migration-CheckerboardSweepCPM.toml
(64^2 periodic, one cell, volume + adhesion, T = 5, Float32; 500 burn-in MCS, then 2,000 MCS
observed every 5 MCS) run on the GPU, 256 in a ReplicaBatch without HIP graphs and 256 in one
replaying 5-MCS graphs. Each replica's
observables must equal what the reference's Julia CPU run archived: the terminal and mean
energy per site, the mean boundary measure and the mixing time exactly, the displacement
statistics to 1e-9. For this workload (one cell in medium, identical von Neumann contact and
surface relations) the global energy is (V - 25)^2 + 12 S with V and S the cell's volume and
boundary measure.
"""
import numpy as np
import pytest

from corepotts import ReplicaBatch, workloads
from corepotts.spec import ALG_CHECKERBOARD
from helpers import (EXACT_MIGRATION_FIELDS, FLOAT_MIGRATION_FIELDS, cell_center, load_migration,
                     migration_summary)


@pytest.mark.gpu
@pytest.mark.parametrize("graph_chunk", [0, 5])
def test_hip_checkerboard_reproduces_realistic_trajectories(graph_chunk):
    col = load_migration()["algorithms"]["CheckerboardSweepCPM"]
    idx = list(range(0 if graph_chunk == 0 else 1, 512, 2))
    batch = ReplicaBatch(ALG_CHECKERBOARD, graph_chunk=graph_chunk)
    slots = [batch.add_model(*workloads.single_cell_migration("checkerboard", int(col["dynamics_seed_hex"][i], 16)),
                             target_mcs=2500) for i in idx]
    series = {s: ([], [], []) for s in slots}
    batch.step(500)
    for _ in range(400):
        batch.step(5)
        for s in slots:
            st = batch.download(s)
            v, b = int(st["finite_volumes"][0]), int(st["boundary_measures"][0])
            words = np.where(st["tags"] == 1, st["ids"].astype(np.int64), -st["ids"].astype(np.int64))
            e, bs, cs = series[s]
            e.append(((v - 25) ** 2 + 12 * b) / 4096)
            bs.append(float(b))
            cs.append(cell_center(words, (64, 64)))
    assert set(batch.progress().values()) == {2500}
    mismatches = []
    for i, s in zip(idx, slots):
        obs = migration_summary(*series[s])
        for key in EXACT_MIGRATION_FIELDS:
            if obs[key] != col[key][i]:
                mismatches.append((i, key, obs[key], col[key][i]))
        for key in FLOAT_MIGRATION_FIELDS:
            if not np.isclose(obs[key], col[key][i], rtol=1e-9, atol=1e-12):
                mismatches.append((i, key, obs[key], col[key][i]))
    batch.close()
    assert not mismatches, f"{len(mismatches)} mismatches, first {mismatches[:5]}"
