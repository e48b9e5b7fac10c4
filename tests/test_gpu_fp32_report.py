"""Float32 configuration report for the 256^3 cube workload (32,768 cubes of 8^3, Float32 J in
[2, 20], 26-neighbour contact, von Neumann boundary, lambda_V = 1.5, lambda_S = 0.5, T = 20):
the accept-decision mismatch rate against the oracle, two-sample Kolmogorov-Smirnov statistics
on energy, cell volumes and boundary lengths, and FP64 energy totals for drift.

* Decisions. Under CheckerboardSweepCPM every mutable site is the recipient of exactly one
  attempt per MCS and a copy changes its recipient only, so from a common state the sites whose
  owner changed in one MCS are exactly the accepted copies; a site whose owner after the MCS
  differs between the GPU and the oracle is a decision made differently. Each MCS starts both
  from the same state (after a mismatch the GPU takes the oracle's state), and the rate is
  mismatches over the decisions the report counts (constraint rejections, acceptance
  rejections and accepted copies). Tracker arrays are compared as well.
* Distributions. Per-cell volumes and boundary measures after the run, GPU against the oracle at
  the same seed (identical samples: D = 0) and the GPU at another dynamics seed against the
  oracle (the check a float configuration needs when decisions differ: p > 0.01). Total energy
  is one number per run, so its KS test runs over an ensemble: 24 seeds on the GPU against 24
  other seeds on the oracle, on the same model at 64^3 (512 cubes).
* Energy. cpm_total_energy and orc_total_energy of the final states (FP64), the FP64 sums of
  accepted delta_h on both sides, and the drift of each: sum - (E(N) - E(0)).

The measured figures go to CPM_REPORT_DIR/fp32_config_report.json when that variable is set.
The MCS count of the full-size lock-step run is CPM_FP32_REPORT_MCS (default 20); the digest
at MCS 200 is compared in test_gpu_deep.py.
"""
import json
import os

import numpy as np
import pytest

from corepotts import ensemble_seed, init_scientific, workloads
from corepotts.native import snapshot
from helpers import oracle_for

pytestmark = pytest.mark.gpu

stats = pytest.importorskip("scipy.stats")

MCS = int(os.environ.get("CPM_FP32_REPORT_MCS", "20"))
TRACKERS = ("finite_volumes", "medium_volumes", "boundary_measures")


def _owner_mismatches(a: dict, b: dict) -> int:
    return int(np.count_nonzero((a["tags"] != b["tags"]) | (a["ids"] != b["ids"])))


def _ks(x, y):
    r = stats.ks_2samp(np.asarray(x, np.float64), np.asarray(y, np.float64))
    return {"D": float(r.statistic), "p": float(r.pvalue), "n": [len(x), len(y)]}


def _write(report):
    out_dir = os.environ.get("CPM_REPORT_DIR")
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        with open(os.path.join(out_dir, "fp32_config_report.json"), "w") as handle:
            json.dump(report, handle, indent=1)
    print("fp32 config report:", json.dumps(report))


_REPORT = {}


def test_fp32_config_decisions_trackers_and_energy(gpu_lib):
    model = workloads.cubes_3d(256)
    compiled = model[0]
    sim, spec = oracle_for(*model)
    gpu = init_scientific(*model)
    e0_gpu, e0_orc = gpu.total_energy(), sim.total_energy()
    decisions = accepted = mismatches = tracker_mismatches = resyncs = 0
    sum_gpu = 0.0
    for k in range(1, MCS + 1):
        gpu.step(1)
        sim.step(1)
        g, o = gpu.download(), sim.store()
        rep = sim.report()
        decisions += rep[9] + rep[10] + rep[11]
        accepted += rep[11]
        sum_gpu += gpu.accepted_energy()
        m = _owner_mismatches(g, o)
        t = sum(int(np.count_nonzero(g[f] != o[f])) for f in TRACKERS)
        mismatches += m
        tracker_mismatches += t
        if m or t:  # continue both from the oracle's state
            resyncs += 1
            gpu.upload(snapshot(compiled, o))
            gpu.mcs = k
    g, o = gpu.download(), sim.store()
    e_gpu, e_orc = gpu.total_energy(), sim.total_energy()
    _, sum_orc = sim.accepted_energy()
    cells = o["active"].astype(bool)
    report = {
        "workload": "256^3 periodic, 32768 cubes of 8^3, Float32 J U[2,20], 26-neighbour contact, "
                    "VN boundary, CheckerboardSweepCPM, seed 1234",
        "mcs": MCS, "decisions": decisions, "accepted": accepted, "mismatches": mismatches,
        "mismatch_rate": mismatches / max(decisions, 1), "tracker_mismatches": tracker_mismatches,
        "resyncs": resyncs,
        "energy_fp64": {"gpu": {"E0": e0_gpu, "EN": e_gpu, "accepted_sum": sum_gpu,
                                "drift": sum_gpu - (e_gpu - e0_gpu)},
                        "oracle": {"E0": e0_orc, "EN": e_orc, "accepted_sum": sum_orc,
                                   "drift": sum_orc - (e_orc - e0_orc)}},
        "ks_same_seed": {"volume": _ks(g["finite_volumes"][cells], o["finite_volumes"][cells]),
                         "boundary": _ks(g["boundary_measures"][cells], o["boundary_measures"][cells])},
    }
    # the GPU under another dynamics seed against the oracle
    other = init_scientific(*model[:4], seed=ensemble_seed(1234, 2))
    other.step(MCS)
    s = other.download()
    report["ks_other_seed"] = {"volume": _ks(s["finite_volumes"][cells], o["finite_volumes"][cells]),
                               "boundary": _ks(s["boundary_measures"][cells], o["boundary_measures"][cells])}
    other.close()
    gpu.close()
    _REPORT.update(report)
    _write(_REPORT)

    assert decisions > 0
    assert mismatches == 0 and tracker_mismatches == 0
    # FP64 sums of Float32 terms in different orders: equal to 1e-12 relative
    assert e0_gpu == pytest.approx(e0_orc, rel=1e-12) and e_gpu == pytest.approx(e_orc, rel=1e-12)
    assert sum_gpu == pytest.approx(sum_orc, rel=1e-12, abs=1e-6)
    # Float32 rounding of each delta_h only: far below one J per hundred copies
    for side in report["energy_fp64"].values():
        assert abs(side["drift"]) < 0.01 * accepted
    assert report["ks_same_seed"]["volume"]["D"] == 0.0
    assert report["ks_same_seed"]["boundary"]["D"] == 0.0
    assert report["ks_other_seed"]["volume"]["p"] > 0.01
    assert report["ks_other_seed"]["boundary"]["p"] > 0.01


def test_fp32_config_energy_distribution_over_seeds(gpu_lib):
    """Total energy after 20 MCS of the 64^3 cube model: 24 GPU runs against 24 oracle runs at
    other seeds, KS p > 0.01; each GPU run also equals the oracle at its own seed."""
    n, mcs = 24, 20
    e_gpu, e_orc = [], []
    for r in range(n):
        gseed, oseed = ensemble_seed(4321, r + 1), ensemble_seed(4321, n + r + 1)
        gpu = init_scientific(*workloads.cubes_3d(64, seed=gseed))
        gpu.step(mcs)
        e_gpu.append(gpu.total_energy())
        gpu.close()
        sim, _ = oracle_for(*workloads.cubes_3d(64, seed=oseed))
        sim.step(mcs)
        e_orc.append(sim.total_energy())
        if r == 0:  # one cross-check at equal seeds
            same, _ = oracle_for(*workloads.cubes_3d(64, seed=gseed))
            same.step(mcs)
            assert same.total_energy() == pytest.approx(e_gpu[0], rel=1e-12)
    ks = _ks(e_gpu, e_orc)
    _REPORT["ks_energy_over_seeds"] = dict(ks, model="64^3 cubes, same parameters", mcs=mcs,
                                           mean_gpu=float(np.mean(e_gpu)), mean_oracle=float(np.mean(e_orc)))
    _write(_REPORT)
    assert ks["p"] > 0.01
