"""corepotts: MI355X-native engine for CorePotts' data-parallel Metropolis sweeps.

Host-side mirror of the CorePotts API surface for CheckerboardSweepCPM and LotteryCPM,
backed by hand-written HIP kernels for gfx950 in libcpm_mi355x.so (C-ABI in
include/cpm_mi355x.h).
"""
from .model import (AlgorithmTemperatureNoise, ConstitutiveMeanInitialization, FixedMechanicalNoise,
                    FluctuatingSurfaceTension, FluctuatingVolumePressure, PreserveMechanicalInitialization,
                    StationaryMechanicalInitialization, TiledCheckerboardCPM, AxisBoundary, BoundaryEdgeCount, BoundaryMeasureTracker, CartesianDomain,
                    CellOwner, CheckerboardSweepCPM, ClosedBoundary, CompiledScientificState,
                    ContactRole, FixedExterior, LogicalPottsState, LotteryCPM, MediumOwner,
                    MediumTypeTable, PeriodicBoundary, ProposalRole, QuadraticBoundaryHamiltonian,
                    QuadraticVolumeHamiltonian, ScientificComponentSet, SequentialCPM,
                    StaticCartesianRelation, SurfaceRole, UnorderedContactHamiltonian,
                    compile_scientific_state, first_shell_relation, moore_offsets,
                    moore_relation, owner_tags_ids, owner_words, rebuild_boundary_measures,
                    required_properties, static_relation, tracker_conformance_errors,
                    validate_relation_domain, von_neumann_offsets)
from .spec import ModelSpec, build_spec, state_arrays
from .native import (DeviceSnapshot, EngineError, ScientificMCSReport, ScientificPottsIntegrator, copy_probe,
                     ensemble_seed, init_scientific, load_library)
from .batch import ReplicaBatch, replica_bytes
from .serve import Completion, JobServer, Request, latency_summary, make_request
from .autotune import BatchShapes, set_row_kernels, set_tiled_staging, set_tile_shape, tile_shapes, tune_tile_shape
from .checkpoint import (CanonicalCheckpoint, CheckpointCompatibilityError, CheckpointIntegrityError,
                         IncompleteCheckpointError, capture_checkpoint, import_checkpoint,
                         load_checkpoint, restore_checkpoint, save_checkpoint, validate_checkpoint)


def step(integrator, n: int = 1):
    """SciMLBase.step!(integrator, n) (sequential.jl:603-609)."""
    return integrator.step(n)


def current_mcs_report(integrator):
    """current_mcs_report(integrator) (sequential.jl:611-613)."""
    return integrator.current_mcs_report()


def logical_state(integrator):
    """logical_state(integrator) (sequential.jl:630-632)."""
    return integrator.logical_state()
