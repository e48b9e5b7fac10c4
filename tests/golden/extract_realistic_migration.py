"""Extracts the reference's realistic-scale single-cell-migration evidence into a fixture.

Source (reference checkout, read as data):
  design/evidence/phase-13/realistic/cpu/single_cell_migration-{SequentialCPM,CheckerboardSweepCPM}.toml
512 replicas per algorithm, produced by integration/transition/RealisticScaleRunner.jl
(run_realistic_replica) with Julia 1.12.6 on the CPU backend, Float32, 1 thread: a 64 x 64
periodic von Neumann lattice holding one cell (a 5 x 5 square at sites 30..34 of each
axis, 1-based), volume (target 25, strength 1) + adhesion (medium-cell 12, cell-cell 4), T = 5,
500 burn-in MCS then 2000 MCS observed every 5 MCS.

Writes tests/golden/realistic_single_cell_migration.json: per algorithm, one column per field
over the 512 replicas (seeds and the observables a trajectory fixes: terminal and mean energy
per site, mean boundary measure, mixing time, centre-of-mass displacement statistics), and the
fields all replicas share (run lengths, the initial-state and model fingerprints).

Usage: python tests/golden/extract_realistic_migration.py [reference root]
"""
import json
import os
import sys

import tomli

# per-replica columns, and fields every replica shares (checked equal, stored once)
FIELDS = ("seed_hex", "dynamics_seed_hex", "energy_per_mutable_site_terminal",
          "energy_per_mutable_site_mean", "boundary_measure_mean", "mixing_time_mcs",
          "net_displacement", "mean_squared_displacement", "speed", "persistence")
SHARED = ("initial_state_fingerprint", "model_fingerprint", "burn_in_mcs", "measured_mcs",
          "observation_cadence_mcs", "volume_coefficient_of_variation_mean")


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    base = os.path.join(root, "design/evidence/phase-13/realistic/cpu")
    out = {"source": "design/evidence/phase-13/realistic/cpu/single_cell_migration-*.toml",
           "algorithms": {}}
    for alg in ("SequentialCPM", "CheckerboardSweepCPM"):
        with open(os.path.join(base, f"single_cell_migration-{alg}.toml"), "rb") as handle:
            doc = tomli.load(handle)
        out["model"] = doc["model"]
        out["environment"] = doc["environment"]
        reps = doc["replica_primary_summaries"]
        shared = {}
        for k in SHARED:
            values = {r[k] for r in reps}
            assert len(values) == 1, k
            shared[k] = values.pop()
        out["shared"] = shared
        out["algorithms"][alg] = {k: [r[k] for r in reps] for k in FIELDS}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "realistic_single_cell_migration.json")
    with open(dst, "w") as handle:
        json.dump(out, handle, separators=(",", ":"))
    print(dst, {a: len(v["seed_hex"]) for a, v in out["algorithms"].items()})


if __name__ == "__main__":
    main()
