"""Oracle against the reference's archived one-MCS transition counts.

For each archived row, 262,144 replicas with seeds 0x1300000000000000 + r each start from
the source state with mcs 0 and run one MCS; the destination-state histogram must equal
the reference's raw_counts exactly (design/evidence/phase-13/empirical/*/row{3,4,9,10}-*.toml,
identical across the reference's cpu, rocm and metal backends). This pins Philox, the
address packing, the realized colouring and colour order, the claim protocol, the Float32
volume + contact energy change, the acceptance law and the commit.
"""
import pytest

from corepotts import CheckerboardSweepCPM, SequentialCPM
from helpers import load_phase13, oracle_for, phase13_fixture

DATA = load_phase13()
ROWS = [(r["row"], r["algorithm"]) for r in DATA["rows"]]


@pytest.mark.parametrize("row_id,algorithm", ROWS)
def test_transition_counts_match_reference(row_id, algorithm):
    row = next(r for r in DATA["rows"] if r["row"] == row_id and r["algorithm"] == algorithm)
    alg_type = CheckerboardSweepCPM if algorithm == "checkerboard" else SequentialCPM
    compiled, proposal, components, alg, labels = phase13_fixture(row, DATA["model"], alg_type)
    sim, _ = oracle_for(compiled, proposal, components, alg, seed=0)
    counts = sim.sample_transitions(int(row["first_seed_hex"], 16), row["replicas"], labels)
    assert counts.sum() == row["replicas"]
    assert [int(c) for c in counts] == row["raw_counts"]


def test_source_state_ids():
    for row in DATA["rows"]:
        compiled, proposal, components, alg, labels = phase13_fixture(row, DATA["model"], SequentialCPM)
        owners = compiled.owners
        ident, scale = 0, 1
        for o in owners:
            ident += labels.index(int(o)) * scale
            scale *= len(labels)
        assert ident + 1 == row["source_state_id"]
