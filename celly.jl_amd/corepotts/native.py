"""ctypes binding of libcpm_mi355x.so (include/cpm_mi355x.h) and the GPU integrator.

The integrator mirrors CorePotts' ScientificPottsIntegrator surface
(lib/CorePotts/src/algorithms/sequential.jl:216-234, 598-632): ``step`` enqueues MCS
without synchronizing, ``current_mcs_report`` costs one launch and one synchronization,
``logical_state`` downloads the state. There is no CPU fallback: a missing or unloadable
library raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import NamedTuple, Optional

import numpy as np

from .model import (MECH_PRESSURE, MECH_STATE_COLUMN, MECH_TENSION, CompiledScientificState,
                    StaticCartesianRelation)
from .spec import ALG_CHECKERBOARD, ALG_LOTTERY, ALG_TILED, ModelSpec, build_spec, state_arrays

MAX_DIRS = 128
MAX_TYPES = 16
MAX_MEDIA = 16
ABI_VERSION = 1

# status codes of the C-ABI (cpm_mi355x.h)
CPM_OK, CPM_E_INVALID, CPM_E_UNSUPPORTED, CPM_E_DEVICE, CPM_E_NOMEM, CPM_E_STATE = range(6)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libcpm_mi355x.so")


class CpmRelation(C.Structure):
    _fields_ = [("count", C.c_int32), ("offsets", (C.c_int32 * 3) * MAX_DIRS),
                ("weights", C.c_double * MAX_DIRS)]


class CpmMechanicsDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("initialization", C.c_int32), ("instance_id", C.c_uint32),
                ("eta", C.c_float), ("noise_is_temperature", C.c_int32), ("noise", C.c_float)]


class CpmModelDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("algorithm", C.c_int32), ("ndim", C.c_int32),
        ("dims", C.c_int32 * 3), ("face_kind", (C.c_int32 * 2) * 3),
        ("face_medium", (C.c_uint32 * 2) * 3), ("mutable_mask", C.c_void_p),
        ("immutable_tags", C.c_void_p), ("immutable_ids", C.c_void_p),
        ("proposal", CpmRelation), ("contact", CpmRelation), ("surface", CpmRelation),
        ("n_energies", C.c_int32), ("energy_order", C.c_int32 * 3),
        ("volume_f64", C.c_int32), ("contact_f64", C.c_int32), ("boundary_f64", C.c_int32),
        ("n_types", C.c_int32), ("J", C.c_double * (MAX_TYPES * MAX_TYPES)),
        ("n_table", C.c_int32), ("table_owner", C.c_uint32 * MAX_MEDIA),
        ("table_type", C.c_uint32 * MAX_MEDIA), ("n_media", C.c_int32),
        ("medium_ids", C.c_uint32 * MAX_MEDIA), ("temperature", C.c_float),
        ("capacity", C.c_uint32), ("seed", C.c_uint64),
        ("retire_target_volume", C.c_double), ("retire_volume_strength", C.c_double),
        ("retire_target_boundary", C.c_int64), ("retire_boundary_strength", C.c_double),
        ("n_mechanics", C.c_int32), ("mechanics", CpmMechanicsDesc * 2),
        ("tile_size", C.c_int32 * 3), ("switching_interval", C.c_int32),
    ]


EXPORTS = (
    "cpm_last_error", "cpm_abi_version", "cpm_create", "cpm_destroy", "cpm_upload_state",
    "cpm_download_state", "cpm_set_seed", "cpm_set_mcs", "cpm_get_mcs", "cpm_step", "cpm_report",
    "cpm_sync", "cpm_schedule", "cpm_site_colors", "cpm_metrics", "cpm_step_timed",
    "cpm_set_graph_chunk", "cpm_prepare_graph", "cpm_copy_probe", "cpm_tile_shape_info", "cpm_set_tile_shape",
    "cpm_set_row_kernels", "cpm_set_tiled_staging", "cpm_kernel_config", "cpm_upload_mechanics", "cpm_download_mechanics",
    "cpm_initialize_mechanics", "cpm_batch_upload_mechanics", "cpm_batch_download_mechanics",
    "cpm_batch_initialize_mechanics", "cpm_tiled_layout", "cpm_model_colors",
    "cpm_ensemble_seed", "cpm_batch_create", "cpm_batch_destroy", "cpm_batch_last_error",
    "cpm_replica_bytes", "cpm_batch_add", "cpm_batch_remove", "cpm_batch_upload_state",
    "cpm_batch_download_state", "cpm_batch_set_target", "cpm_batch_set_mcs", "cpm_batch_step", "cpm_batch_progress",
    "cpm_batch_capacity", "cpm_batch_report", "cpm_batch_sync", "cpm_batch_set_graph_chunk",
    "cpm_batch_bytes", "cpm_report_energy", "cpm_total_energy", "cpm_total_energy_noalloc",
    "cpm_snapshot_capture", "cpm_batch_snapshot_capture", "cpm_snapshot_restore", "cpm_batch_snapshot_restore",
    "cpm_snapshot_download", "cpm_snapshot_download_mechanics", "cpm_snapshot_info", "cpm_snapshot_destroy",
)

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the engine; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libcpm_mi355x.so is not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    P = C.c_void_p
    lib.cpm_last_error.restype = C.c_char_p
    lib.cpm_last_error.argtypes = [P]
    lib.cpm_batch_last_error.restype = C.c_char_p
    lib.cpm_batch_last_error.argtypes = [P]
    lib.cpm_abi_version.restype = C.c_int32
    lib.cpm_create.argtypes = [C.POINTER(CpmModelDesc), C.POINTER(P)]
    lib.cpm_destroy.argtypes = [P]
    state_args = [P] * 11
    lib.cpm_upload_state.argtypes = [P] + state_args
    lib.cpm_download_state.argtypes = [P] + state_args
    lib.cpm_set_seed.argtypes = [P, C.c_uint64]
    lib.cpm_set_mcs.argtypes = [P, C.c_uint64]
    lib.cpm_get_mcs.argtypes = [P, C.POINTER(C.c_uint64)]
    lib.cpm_step.argtypes = [P, C.c_uint64]
    lib.cpm_report.argtypes = [P, C.POINTER(C.c_uint64)]
    lib.cpm_sync.argtypes = [P]
    lib.cpm_report_energy.argtypes = [P, C.POINTER(C.c_double)]
    lib.cpm_total_energy.argtypes = [P, C.POINTER(C.c_double)]
    lib.cpm_total_energy_noalloc.argtypes = [P, C.POINTER(C.c_double)]
    lib.cpm_schedule.argtypes = [P, C.POINTER(C.c_int32), P]
    lib.cpm_site_colors.argtypes = [P, P]
    lib.cpm_metrics.argtypes = [P, C.POINTER(C.c_uint64)]
    lib.cpm_set_graph_chunk.argtypes = [P, C.c_uint32]
    lib.cpm_prepare_graph.argtypes = [P, C.POINTER(C.c_int32)]
    lib.cpm_step_timed.argtypes = [P, C.c_uint64, C.POINTER(C.c_double)]
    lib.cpm_copy_probe.argtypes = [C.c_uint64, C.c_int32, C.POINTER(C.c_double)]
    lib.cpm_tile_shape_info.argtypes = [C.c_int32, C.POINTER(C.c_int32)]
    lib.cpm_set_tile_shape.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    lib.cpm_set_row_kernels.argtypes = [C.c_int32]
    lib.cpm_set_tiled_staging.argtypes = [C.c_int32]
    lib.cpm_kernel_config.argtypes = [P, C.POINTER(C.c_int32)]
    lib.cpm_ensemble_seed.restype = C.c_uint64
    lib.cpm_ensemble_seed.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
    lib.cpm_batch_create.argtypes = [C.c_int32, C.c_uint64, C.POINTER(P)]
    lib.cpm_batch_destroy.argtypes = [P]
    lib.cpm_replica_bytes.argtypes = [C.POINTER(CpmModelDesc), C.POINTER(C.c_uint64)]
    lib.cpm_model_colors.argtypes = [C.POINTER(CpmModelDesc), P, C.POINTER(C.c_int32)]
    lib.cpm_tiled_layout.argtypes = [C.POINTER(CpmModelDesc), P, P, C.POINTER(C.c_int32), P,
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32)]
    lib.cpm_batch_add.argtypes = [P, C.POINTER(CpmModelDesc), C.POINTER(C.c_int32)]
    lib.cpm_batch_remove.argtypes = [P, C.c_int32]
    lib.cpm_batch_upload_state.argtypes = [P, C.c_int32] + state_args
    lib.cpm_batch_download_state.argtypes = [P, C.c_int32] + state_args
    lib.cpm_batch_set_target.argtypes = [P, C.c_int32, C.c_uint64]
    lib.cpm_snapshot_capture.argtypes = [P, C.POINTER(P)]
    lib.cpm_batch_snapshot_capture.argtypes = [P, C.c_int32, C.POINTER(P)]
    lib.cpm_snapshot_restore.argtypes = [P, P]
    lib.cpm_batch_snapshot_restore.argtypes = [P, C.c_int32, P]
    lib.cpm_snapshot_download.argtypes = [P] + state_args
    lib.cpm_snapshot_download_mechanics.argtypes = [P, P, P]
    lib.cpm_snapshot_info.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.cpm_snapshot_destroy.argtypes = [P]
    lib.cpm_batch_set_mcs.argtypes = [P, C.c_int32, C.c_uint64]
    lib.cpm_upload_mechanics.argtypes = [P, P, P]
    lib.cpm_download_mechanics.argtypes = [P, P, P]
    lib.cpm_initialize_mechanics.argtypes = [P]
    lib.cpm_batch_upload_mechanics.argtypes = [P, C.c_int32, P, P]
    lib.cpm_batch_download_mechanics.argtypes = [P, C.c_int32, P, P]
    lib.cpm_batch_initialize_mechanics.argtypes = [P, C.c_int32]
    lib.cpm_batch_step.argtypes = [P, C.c_uint64]
    lib.cpm_batch_progress.argtypes = [P, P, C.c_int32]
    lib.cpm_batch_capacity.argtypes = [P]
    lib.cpm_batch_report.argtypes = [P, C.c_int32, C.POINTER(C.c_uint64)]
    lib.cpm_batch_sync.argtypes = [P]
    lib.cpm_batch_set_graph_chunk.argtypes = [P, C.c_uint32]
    lib.cpm_batch_bytes.argtypes = [P, C.POINTER(C.c_uint64)]
    for name in EXPORTS:
        getattr(lib, name)
    if lib.cpm_abi_version() != ABI_VERSION:
        raise RuntimeError("libcpm_mi355x.so ABI version mismatch")
    _lib = lib
    return lib


class EngineError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"cpm error {code}: {message}")
        self.code = code


def _fill_relation(dst: CpmRelation, rel: Optional[StaticCartesianRelation], ndim: int):
    if rel is None:
        dst.count = 0
        return
    dst.count = len(rel.offsets)
    for k, (o, w) in enumerate(zip(rel.offsets, rel.weights)):
        for a in range(3):
            dst.offsets[k][a] = o[a] if a < ndim else 0
        dst.weights[k] = w


def make_desc(spec: ModelSpec, keep: list) -> CpmModelDesc:
    """Fill cpm_model_desc; arrays referenced by pointer are appended to `keep`."""
    d = CpmModelDesc()
    d.abi_version = ABI_VERSION
    d.algorithm = spec.algorithm
    d.ndim = spec.ndim
    for a in range(3):
        d.dims[a] = spec.dims[a]
        for f in range(2):
            d.face_kind[a][f] = spec.face_kind[a][f]
            d.face_medium[a][f] = spec.face_medium[a][f]
    if spec.mutable_mask is not None:
        arrs = [np.ascontiguousarray(spec.mutable_mask, dtype=np.uint8),
                np.ascontiguousarray(spec.immutable_tags, dtype=np.uint8),
                np.ascontiguousarray(spec.immutable_ids, dtype=np.uint32)]
        keep.extend(arrs)
        d.mutable_mask = arrs[0].ctypes.data
        d.immutable_tags = arrs[1].ctypes.data
        d.immutable_ids = arrs[2].ctypes.data
    _fill_relation(d.proposal, spec.proposal, spec.ndim)
    _fill_relation(d.contact, spec.contact, spec.ndim)
    _fill_relation(d.surface, spec.surface, spec.ndim)
    d.n_energies = len(spec.energy_order)
    for t, term in enumerate(spec.energy_order):
        d.energy_order[t] = term
    d.volume_f64 = int(spec.volume_f64)
    d.contact_f64 = int(spec.contact_f64)
    d.boundary_f64 = int(spec.boundary_f64)
    nt = spec.n_types
    d.n_types = nt
    for i in range(nt):
        for j in range(nt):
            d.J[i * nt + j] = float(spec.J[i, j])
    d.n_table = len(spec.table_owner)
    for i, (o, t) in enumerate(zip(spec.table_owner, spec.table_type)):
        d.table_owner[i] = o
        d.table_type[i] = t
    d.n_media = len(spec.medium_ids)
    for i, m in enumerate(spec.medium_ids):
        d.medium_ids[i] = m
    d.temperature = float(spec.temperature)
    d.capacity = spec.capacity
    d.seed = spec.seed
    d.retire_target_volume, d.retire_volume_strength = spec.retire[0], spec.retire[1]
    d.retire_target_boundary, d.retire_boundary_strength = spec.retire[2], spec.retire[3]
    d.n_mechanics = len(spec.mechanics)
    for k, (kind, init, instance, eta, noise_is_temperature, noise) in enumerate(spec.mechanics):
        m = d.mechanics[k]
        m.kind, m.initialization, m.instance_id = kind, init, instance
        m.eta, m.noise_is_temperature, m.noise = float(eta), int(noise_is_temperature), float(noise)
    for a in range(3):
        d.tile_size[a] = int(spec.tile_size[a])
    d.switching_interval = int(spec.switching_interval)
    return d


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


STATE_FIELDS = ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures", "active",
                "cell_types", "target_volume", "volume_strength", "target_boundary",
                "boundary_strength")


def state_pointers(a: dict) -> list:
    """Pointers of the eleven state arrays in cpm_upload_state / cpm_download_state order."""
    return [_ptr(a.get(k)) for k in STATE_FIELDS]


def state_buffers(spec: ModelSpec) -> dict:
    """Host buffers sized for a download of one replica's state."""
    n = int(np.prod(spec.dims))
    cap = spec.capacity
    return dict(
        tags=np.zeros(n, np.uint8), ids=np.zeros(n, np.uint32),
        finite_volumes=np.zeros(cap, np.int32),
        medium_volumes=np.zeros(max(len(spec.medium_ids), 1), np.int32),
        boundary_measures=np.zeros(cap, np.int64), active=np.zeros(cap, np.uint8),
        cell_types=np.zeros(cap, np.uint32), target_volume=np.zeros(cap, np.float64),
        volume_strength=np.zeros(cap, np.float64), target_boundary=np.zeros(cap, np.int64),
        boundary_strength=np.zeros(cap, np.float64))


def mechanics_columns(spec: ModelSpec):
    """Property keys of the model's mechanical state columns, in (pressure, tension) slot order."""
    kinds = {m[0] for m in spec.mechanics}
    return tuple(MECH_STATE_COLUMN[k] if k in kinds else None for k in (MECH_PRESSURE, MECH_TENSION))


def mechanics_arrays(compiled: CompiledScientificState, spec: ModelSpec):
    """Float32 mechanical state columns for cpm_upload_mechanics (None where absent)."""
    return tuple(None if key is None else np.ascontiguousarray(compiled.properties[key], dtype=np.float32)
                 for key in mechanics_columns(spec))


def snapshot(compiled: CompiledScientificState, s: dict) -> CompiledScientificState:
    """A compiled host state holding the downloaded arrays `s`."""
    snap = compiled.copy()
    snap.tags, snap.ids = s["tags"], s["ids"]
    snap.finite_volumes = s["finite_volumes"]
    snap.medium_volumes = s["medium_volumes"]
    snap.boundary_measures = s["boundary_measures"]
    snap.active = s["active"]
    snap.cell_types = s["cell_types"]
    for name in ("target_volume", "volume_strength", "target_boundary", "boundary_strength",
                 "volume_pressure", "surface_tension"):
        if name in snap.properties and name in s:
            snap.properties[name] = s[name].astype(snap.properties[name].dtype)
    return snap


class DeviceSnapshot:
    """One replica's state at a completed MCS held in device memory (cpm_snapshot_*): owner
    words, cell records, medium volumes, MCS counter and seed. restore() on an integrator or a
    batch slot continues exactly; download() gives the state arrays of a download;
    to_checkpoint() builds the canonical checkpoint (persistence/scientific.jl:250-292) on the
    host from them."""

    def __init__(self, lib, handle, spec: ModelSpec):
        self.lib = lib
        self._s = handle
        self.spec = spec
        mcs, seed, nbytes = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        lib.cpm_snapshot_info(handle, C.byref(mcs), C.byref(seed), C.byref(nbytes))
        self.mcs, self.seed, self.bytes = int(mcs.value), int(seed.value), int(nbytes.value)

    def _check(self, rc: int):
        if rc != 0:
            raise EngineError(rc, self.lib.cpm_last_error(None).decode())

    @property
    def handle(self):
        if not self._s:
            raise RuntimeError("the snapshot has been closed")
        return self._s

    def download(self) -> dict:
        out = state_buffers(self.spec)
        self._check(self.lib.cpm_snapshot_download(self.handle, *state_pointers(out)))
        out["medium_volumes"] = out["medium_volumes"][:len(self.spec.medium_ids)]
        if self.spec.mechanics:
            qp, qt = (None if key is None else np.zeros(self.spec.capacity, np.float32)
                      for key in mechanics_columns(self.spec))
            self._check(self.lib.cpm_snapshot_download_mechanics(self.handle, _ptr(qp), _ptr(qt)))
            for key, arr in zip(mechanics_columns(self.spec), (qp, qt)):
                if key is not None:
                    out[key] = arr
        return out

    def to_checkpoint(self, compiled: CompiledScientificState, ancestry=None):
        from .checkpoint import capture_arrays
        return capture_arrays(self.download(), compiled, self.spec, self.mcs, self.seed, ancestry)

    def close(self):
        if self._s:
            self.lib.cpm_snapshot_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ScientificMCSReport(NamedTuple):
    """Accounting of one completed MCS (sequential.jl:236-251)."""
    mcs: int
    internal_rounds: int
    scheduler_candidates: int
    activated_attempts: int
    realized_proposals: int
    same_owner_no_ops: int
    boundary_no_ops: int
    immutable_recipient_no_ops: int
    dynamic_conflicts: int
    constraint_rejections: int
    acceptance_rejections: int
    accepted_copies: int


class ScientificPottsIntegrator:
    """Device-resident integrator over one lattice, backed by a cpm_ctx."""

    def __init__(self, compiled: CompiledScientificState, spec: ModelSpec):
        if spec.algorithm not in (ALG_CHECKERBOARD, ALG_LOTTERY, ALG_TILED):
            raise NotImplementedError("the MI355X engine runs CheckerboardSweepCPM, LotteryCPM and "
                                      "TiledCheckerboardCPM")
        self.lib = load_library()
        self.spec = spec
        self.compiled = compiled
        keep: list = []
        desc = make_desc(spec, keep)
        handle = C.c_void_p()
        rc = self.lib.cpm_create(C.byref(desc), C.byref(handle))
        if rc != 0:
            raise EngineError(rc, self.lib.cpm_last_error(None).decode())
        self._ctx = handle
        self.upload(compiled)
        self.seed = spec.seed

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None):
            self.lib.cpm_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            raise EngineError(rc, self.lib.cpm_last_error(self._ctx).decode())

    # -- state -------------------------------------------------------------
    def upload(self, compiled: CompiledScientificState):
        a = state_arrays(compiled, self.spec)
        self._check(self.lib.cpm_upload_state(self._ctx, *state_pointers(a)))
        if self.spec.mechanics:
            qp, qt = mechanics_arrays(compiled, self.spec)
            self._check(self.lib.cpm_upload_mechanics(self._ctx, _ptr(qp), _ptr(qt)))

    def initialize_mechanics(self):
        """_initialize_mechanics! on the device state (scientific_mechanics.jl:63-103)."""
        self._check(self.lib.cpm_initialize_mechanics(self._ctx))

    def download(self) -> dict:
        out = state_buffers(self.spec)
        self._check(self.lib.cpm_download_state(self._ctx, *state_pointers(out)))
        out["medium_volumes"] = out["medium_volumes"][:len(self.spec.medium_ids)]
        if self.spec.mechanics:
            qp, qt = (None if key is None else np.zeros(self.spec.capacity, np.float32)
                      for key in mechanics_columns(self.spec))
            self._check(self.lib.cpm_download_mechanics(self._ctx, _ptr(qp), _ptr(qt)))
            for key, arr in zip(mechanics_columns(self.spec), (qp, qt)):
                if key is not None:
                    out[key] = arr
        return out

    def logical_state(self) -> CompiledScientificState:
        """Snapshot of the device state as a compiled host state (logical_state)."""
        return snapshot(self.compiled, self.download())

    # -- stepping ----------------------------------------------------------
    @property
    def mcs(self) -> int:
        v = C.c_uint64()
        self._check(self.lib.cpm_get_mcs(self._ctx, C.byref(v)))
        return int(v.value)

    @mcs.setter
    def mcs(self, value: int):
        self._check(self.lib.cpm_set_mcs(self._ctx, int(value)))

    def set_seed(self, seed: int):
        self._check(self.lib.cpm_set_seed(self._ctx, int(seed)))
        self.seed = int(seed)

    def step(self, n: int = 1):
        if n < 0:
            raise ValueError("scientific step count must be non-negative")
        self._check(self.lib.cpm_step(self._ctx, int(n)))
        return self

    def synchronize(self):
        self._check(self.lib.cpm_sync(self._ctx))

    def current_mcs_report(self) -> Optional[ScientificMCSReport]:
        if self.mcs == 0:
            return None
        out = (C.c_uint64 * 12)()
        self._check(self.lib.cpm_report(self._ctx, out))
        return ScientificMCSReport(*[int(v) for v in out])

    def accepted_energy(self) -> float:
        """FP64 sum of the accepted copies' delta_h over the last MCS (cpm_report_energy)."""
        out = C.c_double()
        self._check(self.lib.cpm_report_energy(self._ctx, C.byref(out)))
        return float(out.value)

    def total_energy(self) -> float:
        """Global energy of the current state in FP64 (cpm_total_energy)."""
        out = C.c_double()
        self._check(self.lib.cpm_total_energy(self._ctx, C.byref(out)))
        return float(out.value)

    def total_energy_noalloc(self) -> float:
        """cpm_total_energy through the library's static scratch (no device allocation)."""
        out = C.c_double()
        self._check(self.lib.cpm_total_energy_noalloc(self._ctx, C.byref(out)))
        return float(out.value)

    def schedule(self):
        count = C.c_int32()
        self._check(self.lib.cpm_schedule(self._ctx, C.byref(count), None))
        order = np.zeros(max(count.value, 1), dtype=np.uint32)
        self._check(self.lib.cpm_schedule(self._ctx, C.byref(count), order.ctypes.data))
        return count.value, order[:count.value]

    def kernel_config(self) -> int:
        """Tile shape id | proposal-count kernels << 4 | row propose kernel << 5 | plain lottery kernel << 6;
        256 for TiledCheckerboardCPM."""
        v = C.c_int32()
        self._check(self.lib.cpm_kernel_config(self._ctx, C.byref(v)))
        return int(v.value)

    def site_colors(self) -> np.ndarray:
        out = np.zeros(int(np.prod(self.spec.dims)), dtype=np.uint8)
        self._check(self.lib.cpm_site_colors(self._ctx, out.ctypes.data))
        return out

    def metrics(self) -> dict:
        out = (C.c_uint64 * 6)()
        self._check(self.lib.cpm_metrics(self._ctx, out))
        keys = ("launches", "host_synchronizations", "host_to_device_transfers",
                "device_to_host_transfers", "device_bytes", "graph_launches")
        return dict(zip(keys, [int(v) for v in out]))

    def step_timed(self, n: int) -> dict:
        """Runs n MCS with HIP events around every kernel (synchronizes)."""
        out = (C.c_double * 8)()
        self._check(self.lib.cpm_step_timed(self._ctx, int(n), out))
        keys = ("propose_ms", "propose_launches", "commit_ms", "commit_launches", "begin_ms",
                "begin_launches", "wall_ms", "mcs")
        return dict(zip(keys, [float(v) for v in out]))

    def set_graph_chunk(self, chunk: int):
        self._check(self.lib.cpm_set_graph_chunk(self._ctx, int(chunk)))

    def snapshot(self) -> DeviceSnapshot:
        """Device-resident snapshot of the state at the current MCS (no host transfer)."""
        h = C.c_void_p()
        self._check(self.lib.cpm_snapshot_capture(self._ctx, C.byref(h)))
        return DeviceSnapshot(self.lib, h, self.spec)

    def restore_snapshot(self, snap: DeviceSnapshot):
        """Continue from a device snapshot of this model (state, seed and MCS counter)."""
        self._check(self.lib.cpm_snapshot_restore(self._ctx, snap.handle))
        self.seed = snap.seed
        return self

    def prepare_graph(self) -> bool:
        """Capture the graph the next steps replay without performing an MCS; True when the
        next step of at least one chunk starts with it."""
        ready = C.c_int32(0)
        self._check(self.lib.cpm_prepare_graph(self._ctx, C.byref(ready)))
        return bool(ready.value)


def init_scientific(state: CompiledScientificState, proposal_relation, components, algorithm,
                    seed: int = 0, initialize_mechanics: bool = True) -> ScientificPottsIntegrator:
    """init_scientific(state, proposal_relation, components, algorithm; seed,
    initialize_mechanics) (checkerboard.jl:197-219, lottery.jl:94-115) on the MI355X engine."""
    spec = build_spec(state, proposal_relation, components, algorithm, seed)
    integ = ScientificPottsIntegrator(state, spec)
    if spec.mechanics and initialize_mechanics:
        integ.initialize_mechanics()
    return integ


def ensemble_seed(master_seed: int, trajectory: int, rerun: int = 1) -> int:
    """ensemble_seed(EnsembleSeedDerivationV1(), master, trajectory, rerun)
    (sciml/interface.jl:1009-1026)."""
    if not 1 <= trajectory <= 0xFFFFFFFF:
        raise ValueError("ensemble trajectory identity must lie in 1:typemax(UInt32)")
    if not 1 <= rerun <= 256:
        raise ValueError("ensemble rerun identity exceeds the v1 address domain")
    return int(load_library().cpm_ensemble_seed(int(master_seed), int(trajectory), int(rerun)))


def model_colors(spec: ModelSpec):
    """(colours per site, colour count) of a checkerboard or tiled model as the engine compiles
    it, on the host (greedy realized colouring, checkerboard.jl:57-90)."""
    lib = load_library()
    keep: list = []
    desc = make_desc(spec, keep)
    colors = np.zeros(int(np.prod(spec.dims)), np.uint8)
    count = C.c_int32()
    rc = lib.cpm_model_colors(C.byref(desc), colors.ctypes.data, C.byref(count))
    if rc != 0:
        raise EngineError(rc, lib.cpm_last_error(None).decode())
    return colors, count.value


def tiled_layout(spec: ModelSpec) -> dict:
    """The engine's TiledCheckerboardCPM layout of a model, computed on the host
    (tiled_layout, tiled_checkerboard.jl:166-188; the batch schedule of
    tiled_checkerboard_device.jl:28-121): grid, reach, halo, tile colours (first axis
    fastest), colour count, switching interval, batches per colour and the positions of the
    box a lane stages in LDS (0: the global-memory kernel)."""
    lib = load_library()
    keep: list = []
    desc = make_desc(spec, keep)
    grid = np.zeros(3, np.int32)
    reach = np.zeros(3, np.int32)
    halo, count, interval, batches, staged = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    rc = lib.cpm_tiled_layout(C.byref(desc), grid.ctypes.data, reach.ctypes.data, C.byref(halo), None,
                              C.byref(count), C.byref(interval), C.byref(batches), C.byref(staged))
    if rc != 0:
        raise EngineError(rc, lib.cpm_last_error(None).decode())
    colors = np.zeros(int(np.prod(grid)), np.uint16)
    rc = lib.cpm_tiled_layout(C.byref(desc), None, None, None, colors.ctypes.data, None, None, None, None)
    if rc != 0:
        raise EngineError(rc, lib.cpm_last_error(None).decode())
    return dict(grid=tuple(int(v) for v in grid), reach=tuple(int(v) for v in reach), halo=halo.value,
                colors=colors, color_count=count.value, interval=interval.value, batches=batches.value,
                staged=staged.value)


def copy_probe(nbytes: int, reps: int = 10) -> float:
    """Mean duration in ms of a plain device copy of `nbytes` bytes (calibration kernel)."""
    lib = load_library()
    out = C.c_double()
    rc = lib.cpm_copy_probe(int(nbytes), int(reps), C.byref(out))
    if rc != 0:
        raise EngineError(rc, lib.cpm_last_error(None).decode())
    return out.value
