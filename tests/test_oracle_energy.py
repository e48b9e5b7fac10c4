"""The oracle's local energy change equals the global energy difference.

Restates lib/CorePotts/test/test_scientific_hamiltonians.jl:47-80: for every actionable
proposal of a fixture, the energy change the oracle evaluates (orc_copy_evaluation, the
fold of the terms in tuple order) must equal global_energy(after) - global_energy(before),
term by term and for all terms together, in Float32 and Float64, in 2D and 3D.

The global energies are written here from the reference's definitions, independently of
the oracle's local-change code:
  volume    scientific_components.jl:97-107   sum over active cells of lambda (V - V*)^2
  contact   scientific_components.jl:315-340  sum over mutable sites and realized relation
            directions (a mutable pair once: direction <= its opposite) of
            w * J[type(owner), type(neighbour)] where the owners differ
  boundary  scientific_components.jl:430-448, 525-538  sum over active cells of
            lambda_S (S - S*)^2, S = unlike-owner incidences over the surface relation
Volumes and boundary measures are recounted from ownership (no tracker is trusted), and a
cell whose last site is copied over leaves the sums (retirement). The tolerance is the
reference's isapprox default, rtol = sqrt(eps(T)) with atol = 0.

orc_total_energy (the oracle's FP64 global energy, used by the drift checks) is checked
against the same restatement on every fixture state.
"""

import numpy as np
import pytest

from corepotts import (AxisBoundary, BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain,
                       CellOwner, ClosedBoundary, ContactRole, LogicalPottsState, MediumOwner,
                       MediumTypeTable, PeriodicBoundary, ProposalRole, QuadraticBoundaryHamiltonian,
                       QuadraticVolumeHamiltonian, ScientificComponentSet, SurfaceRole,
                       UnorderedContactHamiltonian, build_spec, compile_scientific_state,
                       first_shell_relation, required_properties, state_arrays)
from corepotts.model import CheckerboardSweepCPM, moore_relation
from oracle.bridge import OracleSim

ACTIONABLE = 1


# ---------------------------------------------------------------------------
# Global energies restated from the reference (float64 accumulation)
# ---------------------------------------------------------------------------
def _neighbours(dims, periodic, offset):
    """(realized, index) of every site's neighbour at `offset` (column-major, x fastest)."""
    n = int(np.prod(dims))
    rem = np.arange(n)
    ok = np.ones(n, bool)
    lin = np.zeros(n, np.int64)
    stride = 1
    for axis, d in enumerate(dims):
        c = rem % d
        rem = rem // d
        v = c + offset[axis]
        if periodic[axis]:
            v = np.mod(v, d)
        else:
            ok &= (v >= 0) & (v < d)
            v = np.clip(v, 0, d - 1)
        lin += v * stride
        stride *= d
    return ok, lin


class Fixture:
    def __init__(self, dims, periodic, owners, capacity, cell_types, medium_type, T, J,
                 tv, vs, tb, bs, terms, contact_kind):
        self.dims, self.periodic, self.T = tuple(dims), tuple(periodic), T
        self.owners = np.asarray(owners, np.int64)
        self.capacity, self.cell_types, self.medium_type = capacity, dict(cell_types), dict(medium_type)
        self.J = np.asarray(J, T)
        self.tv, self.vs, self.tb, self.bs = (np.asarray(a) for a in (tv, vs, tb, bs))
        self.terms, self.contact_kind = terms, contact_kind
        ndim = len(dims)
        spacing = tuple(T(1.0) for _ in dims)
        self.surface = first_shell_relation(SurfaceRole, ndim, spacing=spacing)
        self.contact = (moore_relation(ContactRole, ndim, spacing=spacing) if contact_kind == "moore"
                        else first_shell_relation(ContactRole, ndim, spacing=spacing))
        self.proposal = first_shell_relation(ProposalRole, ndim, spacing=spacing)

    def owner_type(self, w):
        return self.cell_types[int(w)] if w > 0 else self.medium_type[int(-w)]

    def volume_energy(self, owners):
        total = 0.0
        vols = np.bincount(owners[owners > 0], minlength=self.capacity + 1)
        for c in range(1, self.capacity + 1):
            if vols[c] > 0:
                total += float(self.vs[c - 1]) * (float(vols[c]) - float(self.tv[c - 1])) ** 2
        return total

    def contact_energy(self, owners):
        rel = self.contact
        total = 0.0
        for d, (off, w) in enumerate(zip(rel.offsets, rel.weights)):
            if d + 1 > rel.opposite[d]:
                continue  # a mutable pair once (every site is mutable here)
            ok, lin = _neighbours(self.dims, self.periodic, off)
            for s in np.nonzero(ok)[0]:
                a, b = owners[s], owners[lin[s]]
                if a != b:
                    total += float(w) * float(self.J[self.owner_type(a) - 1, self.owner_type(b) - 1])
        return total

    def boundary_measures(self, owners):
        measure = np.zeros(self.capacity + 1, np.int64)
        for off in self.surface.offsets:
            ok, lin = _neighbours(self.dims, self.periodic, off)
            unlike = ok & (owners[lin] != owners)
            cells = owners[unlike & (owners > 0)]
            np.add.at(measure, cells, 1)
        return measure

    def boundary_energy(self, owners):
        total = 0.0
        vols = np.bincount(owners[owners > 0], minlength=self.capacity + 1)
        measure = self.boundary_measures(owners)
        for c in range(1, self.capacity + 1):
            if vols[c] > 0:
                total += float(self.bs[c - 1]) * (float(measure[c]) - float(self.tb[c - 1])) ** 2
        return total

    def energy(self, owners, terms=None):
        terms = self.terms if terms is None else terms
        f = {"volume": self.volume_energy, "contact": self.contact_energy,
             "boundary": self.boundary_energy}
        return sum(f[t](owners) for t in terms)

    # -- the same model through the host API, fed to the oracle --
    def oracle(self, owners=None):
        T = self.T
        owners = self.owners if owners is None else owners
        comps = []
        schema = {}
        for t in self.terms:
            if t == "volume":
                comps.append(QuadraticVolumeHamiltonian(number_type=T))
            elif t == "contact":
                comps.append(UnorderedContactHamiltonian(self.J, MediumTypeTable(self.medium_type),
                                                         self.contact))
            else:
                comps.append(QuadraticBoundaryHamiltonian(BoundaryEdgeCount(), self.surface, number_type=T))
        for c in comps:
            schema.update(required_properties(c))
        media = sorted(self.medium_type)
        state = LogicalPottsState(owners, self.capacity, cell_types=self.cell_types,
                                  medium_domains=media, property_schema=schema, dims=self.dims)
        for name, values in (("target_volume", self.tv), ("volume_strength", self.vs),
                             ("target_boundary", self.tb), ("boundary_strength", self.bs)):
            if name in state.properties:
                state.properties[name][:] = values.astype(state.properties[name].dtype)
        bnd = tuple(AxisBoundary(PeriodicBoundary() if p else ClosedBoundary()) for p in self.periodic)
        domain = CartesianDomain(self.dims, spacing=tuple(T(1.0) for _ in self.dims), boundaries=bnd)
        compiled = compile_scientific_state(state, domain, BoundaryMeasureTracker(BoundaryEdgeCount(), self.surface))
        spec = build_spec(compiled, self.proposal, ScientificComponentSet(energies=tuple(comps)),
                          CheckerboardSweepCPM(np.float32(10.0)), 7)
        return OracleSim(spec, state_arrays(compiled, spec))


def reference_fixture(dims, T, terms, contact_kind="von_neumann"):
    """lib/CorePotts/test/test_scientific_hamiltonians.jl:1-45 (periodic CartesianDomain)."""
    owners = np.full(dims, MediumOwner(1), np.int64)
    if len(dims) == 2:
        owners[0, 0] = owners[1, 0] = owners[0, 1] = CellOwner(1)
        owners[-1, -1] = owners[-2, -1] = owners[-1, -2] = CellOwner(2)
    else:
        owners[0, 0, 0] = owners[1, 0, 0] = owners[0, 1, 0] = CellOwner(1)
        owners[-1, -1, -1] = owners[-2, -1, -1] = owners[-1, -2, -1] = CellOwner(2)
    return Fixture(dims, (True,) * len(dims), owners.reshape(-1, order="F"), 2, {1: 2, 2: 2}, {1: 1}, T,
                   [[0, 3], [3, 1]], [T(4)] * 2, [T(1.25)] * 2, [8, 8], [T(0.75)] * 2, terms, contact_kind)


def random_fixture(seed, dims, periodic, T, terms, contact_kind):
    """Three cells (one of a single site, so retirement is exercised) and two media, Float
    parameters that are not small binary fractions, a random symmetric J over four types."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    owners = rng.choice(np.array([1, 2, 3, -1, -2], np.int64), size=n, p=[0.3, 0.25, 0.0, 0.3, 0.15])
    owners[rng.integers(n)] = 3  # a one-site cell
    upper = rng.uniform(0.0, 10.0, size=(4, 4))
    J = np.triu(upper) + np.triu(upper, 1).T
    return Fixture(dims, periodic, owners, 3, {1: 2, 2: 3, 3: 4}, {1: 1, 2: 2}, T, J.astype(T),
                   rng.uniform(1.0, 6.0, 3).astype(T), rng.uniform(0.5, 2.0, 3).astype(T),
                   rng.integers(2, 12, 3), rng.uniform(0.1, 1.0, 3).astype(T), terms, contact_kind)


TERM_SETS = [("volume",), ("contact",), ("boundary",), ("volume", "contact", "boundary"),
             ("contact", "boundary", "volume")]


def _isapprox(x, y, T, atol=0.0):
    rtol = float(np.sqrt(np.finfo(T).eps))
    return abs(x - y) <= atol + rtol * max(abs(x), abs(y))


def _check_all_proposals(fx: Fixture):
    """Every actionable proposal: local energy change == global difference. A change that
    cancels to zero in exact arithmetic leaves Float32 rounding noise in the local fold, so
    beside the reference's rtol an absolute 64 eps(T) of the fixture's largest change is
    allowed (a wrong, dropped or doubled term moves the value by a whole J, lambda or w)."""
    sim = fx.oracle()
    before = fx.energy(fx.owners)
    ndir = fx.proposal.direction_count()
    checked = retired = 0
    expected_all = {}
    for recipient in range(1, fx.owners.size + 1):
        for direction in range(1, ndir + 1):
            ok, lin = _neighbours(fx.dims, fx.periodic, fx.proposal.offsets[direction - 1])
            if ok[recipient - 1] and fx.owners[lin[recipient - 1]] != fx.owners[recipient - 1]:
                after_owners = fx.owners.copy()
                after_owners[recipient - 1] = fx.owners[lin[recipient - 1]]
                expected_all[(recipient, direction)] = fx.energy(after_owners) - before
    atol = 64 * float(np.finfo(fx.T).eps) * max([abs(v) for v in expected_all.values()] + [1.0])
    for recipient in range(1, fx.owners.size + 1):
        for direction in range(1, ndir + 1):
            outcome, dh32, mech = sim.copy_evaluation(recipient, direction)
            outcome2, dh = sim.copy_energy_fold(recipient, direction)
            assert outcome2 == outcome
            if outcome != ACTIONABLE:
                continue
            assert mech == 0.0
            # evaluate_copy rounds the fold once to the algorithm's Float32
            assert dh32 == np.float32(dh)
            expected = expected_all.pop((recipient, direction))
            assert _isapprox(float(dh), expected, fx.T, atol), \
                (fx.terms, fx.contact_kind, recipient, direction, float(dh), expected)
            checked += 1
            losing = fx.owners[recipient - 1]
            retired += bool(losing > 0 and np.count_nonzero(fx.owners == losing) == 1)
    assert checked > 0
    assert not expected_all  # the oracle found every actionable proposal
    return checked, retired


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("dims", [(4, 4), (4, 4, 4)])
@pytest.mark.parametrize("terms", TERM_SETS)
@pytest.mark.parametrize("contact_kind", ["von_neumann", "moore"])
def test_reference_fixture_local_equals_global(T, dims, terms, contact_kind):
    if contact_kind == "moore" and "contact" not in terms:
        pytest.skip("the contact relation only matters with the contact term")
    _check_all_proposals(reference_fixture(dims, T, terms, contact_kind))


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("geometry", [((5, 6), (True, True)), ((5, 4), (False, True)),
                                      ((3, 4, 5), (True, True, True)), ((4, 3, 4), (True, False, True))])
@pytest.mark.parametrize("terms", TERM_SETS)
def test_random_fixture_local_equals_global(seed, T, geometry, terms):
    dims, periodic = geometry
    for contact_kind in ("von_neumann", "moore"):
        if contact_kind == "moore" and "contact" not in terms:
            continue
        if any(p and d < 3 for p, d in zip(periodic, dims)):
            continue
        checked, retired = _check_all_proposals(random_fixture(seed, dims, periodic, T, terms, contact_kind))
        assert retired > 0  # the one-site cell's last site was copied over at least once


@pytest.mark.parametrize("seed", [4, 5])
@pytest.mark.parametrize("geometry", [((5, 6), (True, True)), ((4, 3, 4), (True, False, True))])
def test_oracle_total_energy_matches_restatement(seed, geometry):
    dims, periodic = geometry
    for T in (np.float32, np.float64):
        for contact_kind in ("von_neumann", "moore"):
            fx = random_fixture(seed, dims, periodic, T, ("volume", "contact", "boundary"), contact_kind)
            sim = fx.oracle()
            assert sim.total_energy() == pytest.approx(fx.energy(fx.owners), rel=1e-12)


def test_energy_term_types_follow_julia_promotion():
    """The arithmetic type of a term is float(promote_type(...)) of its columns and relation
    weights under Julia's rules (scientific_components.jl:101, 529): the Float32 fixture of
    test_scientific_hamiltonians.jl keeps every term Float32 (`global_energy(...) isa T`), even
    though target_boundary is Int64; Float64 relation weights promote the boundary term."""
    from corepotts.spec import julia_float_promotion
    assert julia_float_promotion(np.int64, np.float32, np.float32) == np.float32
    assert julia_float_promotion(np.int64, np.float32, np.float64) == np.float64
    assert julia_float_promotion(np.float32, np.float32) == np.float32
    assert julia_float_promotion(np.float32, np.float64) == np.float64
    assert julia_float_promotion(np.int64) == np.float64
    for T, f64 in ((np.float32, False), (np.float64, True)):
        for dims in ((4, 4), (4, 4, 4)):
            fx = reference_fixture(dims, T, ("volume", "contact", "boundary"))
            spec = fx.oracle().spec
            assert (spec.volume_f64, spec.contact_f64, spec.boundary_f64) == (f64, f64, f64)
