"""The oracle's colour passes and lottery rounds run on OpenMP threads when a colour holds at
least 16,384 sites; claims are order-free maxima and minima and the accepted copies are
committed in site order, so the state, the report and the FP64 energy sums must not depend on
the thread count."""
import numpy as np
import pytest

from corepotts import build_spec, state_arrays, workloads
from oracle.bridge import OracleSim, set_threads
from helpers import state_digest


def _run(model, mcs, threads):
    set_threads(threads)
    spec = build_spec(*model)
    sim = OracleSim(spec, state_arrays(model[0], spec))
    sim.step(mcs)
    return state_digest(sim.store()), tuple(sim.report()), sim.accepted_energy()


@pytest.mark.parametrize("algorithm,mcs", [("checkerboard", 3), ("lottery", 1)])
def test_results_do_not_depend_on_thread_count(algorithm, mcs):
    model = workloads.scaled_2d(512, algorithm=algorithm)  # 65,536 sites per colour
    try:
        one = _run(model, mcs, 1)
        many = _run(model, mcs, 4)
    finally:
        set_threads(0)
    assert one == many
