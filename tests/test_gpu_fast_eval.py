"""The commit kernel's specialised evaluations against the oracle.

Periodic obstacle-free lattices with Float32 energy terms and no mechanics take a commit
evaluation whose loops are fixed at compile time (2D: 0, 4 or 8 contact directions; 3D: 0,
6 or 26). Every kind runs here with non-integer Float32 parameters (contact matrix, volume
and boundary strengths and targets) and each energy order, for CheckerboardSweepCPM and
LotteryCPM, and must stay bit-exact against the oracle; the kernel configuration confirms
which evaluation ran. A model with a Float64 term takes the general evaluation.
"""
import numpy as np
import pytest

from corepotts import (BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain, CheckerboardSweepCPM,
                       ContactRole, LogicalPottsState, LotteryCPM, MediumTypeTable, ProposalRole,
                       QuadraticBoundaryHamiltonian, QuadraticVolumeHamiltonian, ScientificComponentSet,
                       SurfaceRole, UnorderedContactHamiltonian, compile_scientific_state,
                       first_shell_relation, moore_relation, required_properties)
from test_gpu_parity import run_parity

pytestmark = pytest.mark.gpu


def _model(ndim, contact, order, seed, f64_volume=False):
    rng = np.random.default_rng(seed)
    L = 36 if ndim == 2 else 16
    side = 6 if ndim == 2 else 4
    dims = (L,) * ndim
    idx = np.arange(L)
    grids = np.meshgrid(*([idx // side] * ndim), indexing="ij")
    per = L // side
    cell = np.zeros(dims, np.int64)
    stride = 1
    for gr in grids:
        cell += gr * stride
        stride *= per
    cell += 1
    inside = np.logical_and.reduce(np.meshgrid(*([(idx % side) < side - 1] * ndim), indexing="ij"))
    owners = np.where(inside, cell, -1).reshape(-1, order="F")
    ncells = per ** ndim
    types = {k: 2 + (k - 1) % 2 for k in range(1, ncells + 1)}
    sp = (np.float32(1.0),) * ndim  # Float32 spacing: relation weights do not promote a term to Float64
    surface = first_shell_relation(SurfaceRole, ndim, spacing=sp)
    comps = {}
    comps["volume"] = QuadraticVolumeHamiltonian(number_type=np.float64 if f64_volume else np.float32)
    comps["boundary"] = QuadraticBoundaryHamiltonian(BoundaryEdgeCount(), surface, number_type=np.float32)
    if contact:
        rel = (moore_relation(ContactRole, ndim, spacing=sp) if contact == "moore"
               else first_shell_relation(ContactRole, ndim, spacing=sp))
        upper = rng.uniform(1.0, 12.0, size=(3, 3))
        J = (np.triu(upper) + np.triu(upper, 1).T).astype(np.float32)
        comps["contact"] = UnorderedContactHamiltonian(J, MediumTypeTable({1: 1}), rel)
    energies = tuple(comps[t] for t in order if t in comps)
    schema = {}
    for c in energies:
        schema.update(required_properties(c))
    state = LogicalPottsState(owners, ncells, cell_types=types, medium_domains=(1,), property_schema=schema,
                              dims=dims)
    vol = (side - 1) ** ndim
    for name, values in (("target_volume", rng.uniform(0.8, 1.2, ncells) * vol),
                         ("volume_strength", rng.uniform(0.3, 1.7, ncells)),
                         ("target_boundary", rng.integers(2 * ndim * (side - 2) ** (ndim - 1), 2 * ndim * side ** (ndim - 1), ncells)),
                         ("boundary_strength", rng.uniform(0.05, 0.6, ncells))):
        if name in state.properties:
            state.properties[name][:] = values.astype(state.properties[name].dtype)
    compiled = compile_scientific_state(state, CartesianDomain(dims, spacing=sp),
                                        BoundaryMeasureTracker(BoundaryEdgeCount(), surface))
    return compiled, first_shell_relation(ProposalRole, ndim, spacing=sp), ScientificComponentSet(energies=energies)


KINDS = {(2, None): 1, (2, "vn"): 2, (2, "moore"): 3, (3, None): 4, (3, "vn"): 5, (3, "moore"): 6}


@pytest.mark.parametrize("alg_type", [CheckerboardSweepCPM, LotteryCPM])
@pytest.mark.parametrize("ndim,contact", list(KINDS))
@pytest.mark.parametrize("order", [("volume", "contact", "boundary"), ("boundary", "volume", "contact")])
def test_fast_evaluation_bit_exact(alg_type, ndim, contact, order):
    compiled, proposal, components = _model(ndim, contact, order, seed=ndim * 10 + len(str(contact)))
    gpu, sim, reps = run_parity(compiled, proposal, components, alg_type(np.float32(6.0)), 4242, [1, 3, 6])
    assert (gpu.kernel_config() >> 8) & 15 == KINDS[(ndim, contact)]
    assert reps[-1].accepted_copies > 0


def test_float64_term_takes_general_evaluation():
    compiled, proposal, components = _model(2, "moore", ("volume", "contact", "boundary"), seed=5, f64_volume=True)
    gpu, sim, reps = run_parity(compiled, proposal, components, CheckerboardSweepCPM(np.float32(6.0)), 77, [1, 4])
    assert (gpu.kernel_config() >> 8) & 15 == 0
