"""Extract the reference's archived one-MCS transition counts into a JSON fixture.

The reference archives, for four source states of two tiny fixtures, the histogram of
destination states over 262,144 independent seeded replicas of one MCS
(design/evidence/phase-13/empirical/<backend>/row{3,4,9,10}-<algorithm>-qualification.toml,
line 1 ``raw_counts``; sampling section ``first_seed_hex`` .. ``last_seed_hex``). The
histograms are identical across the reference's cpu, rocm and metal backends; this script
checks that and records the fixture parameters from
design/audits/phase-13-fixture-manifest.toml next to the counts.

Run in a container that has the reference checkout at /root/reference:
    python tests/golden/extract_phase13_counts.py
Writes tests/golden/phase13_transition_counts.json (data only: inputs and expected counts).
"""
import json
import os
import sys

import tomli

REF = "/root/reference"
EVIDENCE = os.path.join(REF, "design", "evidence", "phase-13", "empirical")
MANIFEST = os.path.join(REF, "design", "audits", "phase-13-fixture-manifest.toml")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "phase13_transition_counts.json")


def main():
    with open(MANIFEST, "rb") as handle:
        manifest = tomli.load(handle)
    fixtures = {fixture["id"]: fixture for fixture in manifest["fixtures"]}
    rows = []
    for row in (3, 4, 9, 10):
        for algorithm in ("checkerboard", "sequential"):
            name = f"row{row}-{algorithm}-qualification.toml"
            per_backend = {}
            for backend in ("cpu", "rocm", "metal"):
                with open(os.path.join(EVIDENCE, backend, name), "rb") as handle:
                    per_backend[backend] = tomli.load(handle)
            counts = {b: tuple(doc["raw_counts"]) for b, doc in per_backend.items()}
            if len(set(counts.values())) != 1:
                sys.exit(f"{name}: raw_counts differ across backends")
            doc = per_backend["cpu"]
            fixture_id = doc["identity"]["fixture"]
            fixture = fixtures[fixture_id]
            rows.append({
                "row": row,
                "algorithm": algorithm,
                "fixture_id": fixture_id,
                "source_state_encoding": doc["identity"]["source_state_encoding"],
                "source_state_id": doc["sampling"]["source_state_id"],
                "dimensions": fixture["dimensions"],
                "topology": fixture["topology"],
                "boundaries": fixture["boundaries"],
                "temperature": fixture["temperature"],
                "energy_scale": fixture["energy_scale"],
                "first_seed_hex": doc["sampling"]["first_seed_hex"],
                "replicas": doc["sampling"]["replicas"],
                "raw_counts": list(counts["cpu"]),
                "backends_agreeing": sorted(per_backend),
                "source_file": f"design/evidence/phase-13/empirical/cpu/{name}",
            })
    model = manifest["model"]
    payload = {
        "description": "one-MCS destination-state histograms over seeded replicas",
        "model": {
            "medium_owner_id": model["medium_owner_id"],
            "medium_type_id": model["medium_type_id"],
            "cell_type_id": model["cell_type_id"],
            "cell_capacity": model["cell_capacity"],
            "target_volume": model["target_volume"],
            "contact_matrix_base_row_major": model["contact_matrix_base_row_major"],
            "production_real_type": manifest["empirical_sampling"]["production_real_type"],
        },
        "rows": rows,
    }
    with open(OUT, "w") as handle:
        json.dump(payload, handle, indent=1)
    print(f"wrote {OUT} with {len(rows)} rows")


if __name__ == "__main__":
    main()
