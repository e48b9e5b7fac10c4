"""ctypes bridge to the CPU oracle (liboracle_cpm.so). TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the CPU-baseline leg of bench.py use this module; the
product library never loads it. It feeds the oracle the same lowered model description
(corepotts.spec.ModelSpec) that the engine receives.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_cpm.so")
MAX_DIRS, MAX_TYPES, MAX_MEDIA = 128, 16, 16


class OrcRelation(C.Structure):
    _fields_ = [("n", C.c_int32), ("off", (C.c_int32 * 3) * MAX_DIRS),
                ("weight", C.c_double * MAX_DIRS)]


class OrcModel(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32), ("dims", C.c_int32 * 3), ("face_kind", (C.c_int32 * 2) * 3),
        ("face_medium", (C.c_uint32 * 2) * 3), ("mutable_mask", C.c_void_p),
        ("imm_tags", C.c_void_p), ("imm_ids", C.c_void_p),
        ("proposal", OrcRelation), ("contact", OrcRelation), ("surface", OrcRelation),
        ("n_energies", C.c_int32), ("energy_order", C.c_int32 * 3),
        ("volume_f64", C.c_int32), ("contact_f64", C.c_int32), ("boundary_f64", C.c_int32),
        ("n_types", C.c_int32), ("J", C.c_double * (MAX_TYPES * MAX_TYPES)),
        ("n_table", C.c_int32), ("table_owner", C.c_uint32 * MAX_MEDIA),
        ("table_type", C.c_uint32 * MAX_MEDIA), ("n_media", C.c_int32),
        ("medium_ids", C.c_uint32 * MAX_MEDIA), ("temperature", C.c_float),
        ("capacity", C.c_int32), ("retire_target_volume", C.c_double),
        ("retire_volume_strength", C.c_double), ("retire_target_boundary", C.c_int64),
        ("retire_boundary_strength", C.c_double),
        ("n_mech", C.c_int32), ("mech_kind", C.c_int32 * 2), ("mech_init", C.c_int32 * 2),
        ("mech_instance", C.c_uint32 * 2), ("mech_eta", C.c_float * 2),
        ("mech_noise_is_temperature", C.c_int32 * 2), ("mech_noise", C.c_float * 2),
        ("tile_size", C.c_int32 * 3), ("switching_interval", C.c_int32),
    ]


class OrcState(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in (
        "tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures", "active",
        "cell_types", "target_volume", "volume_strength", "target_boundary",
        "boundary_strength", "volume_pressure", "surface_tension")]


_lib = None


def build_oracle() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle_cpm.so"], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build_oracle()
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    U64, U32 = C.c_uint64, C.c_uint32
    lib.orc_philox4x32_10.argtypes = [P, P, P]
    lib.orc_counter_key.argtypes = [U64, U32, U64, U32, U32, U32, U32, U64, U32, U32, P]
    lib.orc_rng_words.argtypes = [U64, U32, U64, U32, U32, U32, U32, U64, U32, U32, P]
    lib.orc_uniform_open01_f32.restype = C.c_float
    lib.orc_uniform_open01_f32.argtypes = [U32]
    lib.orc_bounded_uint.restype = U32
    lib.orc_bounded_uint.argtypes = [U64, U32, U64, U32, U32, U32, U32, U32, U32]
    lib.orc_ensemble_seed.restype = U64
    lib.orc_ensemble_seed.argtypes = [U64, U32, U32]
    lib.orc_realized_coloring.restype = C.c_int32
    lib.orc_realized_coloring.argtypes = [C.POINTER(OrcModel), P, C.POINTER(C.c_int32)]
    lib.orc_conflict_degrees.restype = C.c_int32
    lib.orc_conflict_degrees.argtypes = [C.POINTER(OrcModel), P]
    lib.orc_create.restype = P
    lib.orc_create.argtypes = [C.POINTER(OrcModel), C.c_int32, U64]
    lib.orc_destroy.argtypes = [P]
    lib.orc_state_of.restype = C.POINTER(OrcState)
    lib.orc_state_of.argtypes = [P]
    lib.orc_set_seed.argtypes = [P, U64]
    lib.orc_set_mcs.argtypes = [P, U64]
    lib.orc_mcs.restype = U64
    lib.orc_mcs.argtypes = [P]
    lib.orc_round_count.restype = C.c_int32
    lib.orc_round_count.argtypes = [P]
    lib.orc_step.restype = C.c_int32
    lib.orc_step.argtypes = [P, U64]
    lib.orc_report.restype = C.c_int32
    lib.orc_report.argtypes = [P, P]
    lib.orc_last_order.restype = C.c_int32
    lib.orc_last_order.argtypes = [P, P]
    lib.orc_last_activations.restype = C.c_int64
    lib.orc_last_activations.argtypes = [P, P]
    lib.orc_load_state.argtypes = [P, C.POINTER(OrcState)]
    lib.orc_store_state.argtypes = [P, C.POINTER(OrcState)]
    lib.orc_tracker_errors.restype = C.c_int32
    lib.orc_tracker_errors.argtypes = [P]
    lib.orc_claims.argtypes = [C.c_int32, P, P, P, P, C.c_int32, P, P, P, P]
    lib.orc_sample_transitions.restype = C.c_int32
    lib.orc_sample_transitions.argtypes = [P, U64, U64, P, C.c_int32, P]
    lib.orc_copy_evaluation.restype = C.c_int32
    lib.orc_copy_evaluation.argtypes = [P, U32, C.c_int32, P]
    lib.orc_copy_energy_fold.restype = C.c_int32
    lib.orc_copy_energy_fold.argtypes = [P, U32, C.c_int32, C.POINTER(C.c_double)]
    lib.orc_set_tiled_reconcile.argtypes = [P, C.c_int32]
    lib.orc_tiled_layout.restype = C.c_int32
    lib.orc_tiled_layout.argtypes = [P, P, P, P, P]
    lib.orc_initialize_mechanics.argtypes = [P]
    lib.orc_advance_mechanics.argtypes = [P, U64, U32, U32, C.c_float]
    lib.orc_normal_box_muller.restype = C.c_float
    lib.orc_normal_box_muller.argtypes = [U64, U32, U64, U32, U32, U32, U32]
    lib.orc_ou_transition.restype = C.c_float
    lib.orc_ou_transition.argtypes = [C.c_float] * 7
    lib.orc_total_energy.restype = C.c_double
    lib.orc_total_energy.argtypes = [P]
    lib.orc_accepted_energy.argtypes = [P, P]
    lib.orc_set_threads.argtypes = [C.c_int32]
    lib.orc_threads.restype = C.c_int32
    _lib = lib
    return lib


def _rel(dst: OrcRelation, rel, ndim: int):
    if rel is None:
        dst.n = 0
        return
    dst.n = len(rel.offsets)
    for k, (o, w) in enumerate(zip(rel.offsets, rel.weights)):
        for a in range(3):
            dst.off[k][a] = o[a] if a < ndim else 0
        dst.weight[k] = w


def make_model(spec, keep: list) -> OrcModel:
    m = OrcModel()
    m.ndim = spec.ndim
    for a in range(3):
        m.dims[a] = spec.dims[a]
        for f in range(2):
            m.face_kind[a][f] = spec.face_kind[a][f]
            m.face_medium[a][f] = spec.face_medium[a][f]
    if spec.mutable_mask is not None:
        arrs = [np.ascontiguousarray(spec.mutable_mask, dtype=np.uint8),
                np.ascontiguousarray(spec.immutable_tags, dtype=np.uint8),
                np.ascontiguousarray(spec.immutable_ids, dtype=np.uint32)]
        keep.extend(arrs)
        m.mutable_mask, m.imm_tags, m.imm_ids = (a.ctypes.data for a in arrs)
    _rel(m.proposal, spec.proposal, spec.ndim)
    _rel(m.contact, spec.contact, spec.ndim)
    _rel(m.surface, spec.surface, spec.ndim)
    m.n_energies = len(spec.energy_order)
    for t, term in enumerate(spec.energy_order):
        m.energy_order[t] = term
    m.volume_f64 = int(spec.volume_f64)
    m.contact_f64 = int(spec.contact_f64)
    m.boundary_f64 = int(spec.boundary_f64)
    nt = spec.n_types
    m.n_types = nt
    for i in range(nt):
        for j in range(nt):
            m.J[i * nt + j] = float(spec.J[i, j])
    m.n_table = len(spec.table_owner)
    for i, (o, t) in enumerate(zip(spec.table_owner, spec.table_type)):
        m.table_owner[i] = o
        m.table_type[i] = t
    m.n_media = len(spec.medium_ids)
    for i, mid in enumerate(spec.medium_ids):
        m.medium_ids[i] = mid
    m.temperature = float(spec.temperature)
    m.capacity = spec.capacity
    m.retire_target_volume, m.retire_volume_strength = spec.retire[0], spec.retire[1]
    m.retire_target_boundary, m.retire_boundary_strength = spec.retire[2], spec.retire[3]
    m.n_mech = len(spec.mechanics)
    for k, (kind, init, instance, eta, noise_is_temperature, noise) in enumerate(spec.mechanics):
        m.mech_kind[k], m.mech_init[k], m.mech_instance[k] = kind, init, instance
        m.mech_eta[k], m.mech_noise_is_temperature[k], m.mech_noise[k] = float(eta), int(noise_is_temperature), float(noise)
    for a in range(3):
        m.tile_size[a] = spec.tile_size[a]
    m.switching_interval = spec.switching_interval
    return m


STATE_FIELDS = ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures", "active",
                "cell_types", "target_volume", "volume_strength", "target_boundary",
                "boundary_strength")
MECH_FIELDS = ("volume_pressure", "surface_tension")   # mechanics kinds 1 and 2


class OracleSim:
    """CPU reference integrator over one lattice (sequential / checkerboard / lottery)."""

    def __init__(self, spec, arrays: dict, algorithm: Optional[int] = None):
        self.lib = load()
        self.spec = spec
        self._keep: list = []
        self.model = make_model(spec, self._keep)
        alg = spec.algorithm if algorithm is None else algorithm
        self._sim = self.lib.orc_create(C.byref(self.model), alg, spec.seed)
        if not self._sim:
            raise RuntimeError("oracle rejected the model")
        self.load(arrays)

    def __del__(self):
        try:
            if self._sim:
                self.lib.orc_destroy(self._sim)
                self._sim = None
        except Exception:
            pass

    def _state_struct(self, arrays: dict) -> OrcState:
        s = OrcState()
        for name in STATE_FIELDS + MECH_FIELDS:
            a = arrays.get(name)
            setattr(s, name, None if a is None else a.ctypes.data)
        return s

    def load(self, arrays: dict):
        med = np.zeros(MAX_MEDIA, np.int32)
        med[:len(self.spec.medium_ids)] = arrays["medium_volumes"][:len(self.spec.medium_ids)]
        a = dict(arrays)
        a["medium_volumes"] = med
        a = {k: np.ascontiguousarray(v) for k, v in a.items()}
        self._loaded = a
        self.lib.orc_load_state(self._sim, C.byref(self._state_struct(a)))

    def store(self) -> dict:
        n = int(np.prod(self.spec.dims))
        cap = self.spec.capacity
        out = dict(tags=np.zeros(n, np.uint8), ids=np.zeros(n, np.uint32),
                   finite_volumes=np.zeros(cap, np.int32), medium_volumes=np.zeros(MAX_MEDIA, np.int32),
                   boundary_measures=np.zeros(cap, np.int64), active=np.zeros(cap, np.uint8),
                   cell_types=np.zeros(cap, np.uint32), target_volume=np.zeros(cap, np.float64),
                   volume_strength=np.zeros(cap, np.float64), target_boundary=np.zeros(cap, np.int64),
                   boundary_strength=np.zeros(cap, np.float64))
        for name in self._mech_keys():
            out[name] = np.zeros(cap, np.float32)
        self.lib.orc_store_state(self._sim, C.byref(self._state_struct(out)))
        out["medium_volumes"] = out["medium_volumes"][:len(self.spec.medium_ids)]
        return out

    def _mech_keys(self):
        return [MECH_FIELDS[kind - 1] for kind, *_ in self.spec.mechanics]

    def copy_evaluation(self, recipient: int, direction: int):
        """(outcome, delta_h, mechanical_work) of one proposal."""
        out = np.zeros(2, np.float32)
        rc = self.lib.orc_copy_evaluation(self._sim, int(recipient), int(direction), out.ctypes.data)
        return int(rc), out[0], out[1]

    def copy_energy_fold(self, recipient: int, direction: int):
        """(outcome, energy fold in the terms' own types, before the Float32 rounding)."""
        out = C.c_double()
        rc = self.lib.orc_copy_energy_fold(self._sim, int(recipient), int(direction), C.byref(out))
        return int(rc), out.value

    def set_tiled_reconcile(self, recount: bool):
        """Reconcile tiled subrounds by recount from ownership (True) or snapshot plus deltas."""
        self.lib.orc_set_tiled_reconcile(self._sim, 1 if recount else 0)

    def tiled_layout(self):
        """(grid, reach, halo, tile colours) of a TiledCheckerboardCPM simulation."""
        grid = np.zeros(3, np.int32)
        reach = np.zeros(3, np.int32)
        halo = C.c_int32()
        self.lib.orc_tiled_layout(self._sim, grid.ctypes.data, reach.ctypes.data, C.byref(halo), None)
        colors = np.zeros(int(np.prod(grid)), np.uint16)
        count = self.lib.orc_tiled_layout(self._sim, grid.ctypes.data, reach.ctypes.data, C.byref(halo),
                                          colors.ctypes.data)
        return tuple(int(g) for g in grid), tuple(int(r) for r in reach), int(halo.value), colors, int(count)

    def initialize_mechanics(self):
        self.lib.orc_initialize_mechanics(self._sim)

    def advance_mechanics(self, mcs: int, subround: int, phase: int, interval: float):
        self.lib.orc_advance_mechanics(self._sim, int(mcs), int(subround), int(phase), float(interval))

    def step(self, n: int = 1):
        self.lib.orc_step(self._sim, int(n))

    def report(self):
        out = np.zeros(12, np.uint64)
        if self.lib.orc_report(self._sim, out.ctypes.data) != 0:
            return None
        return tuple(int(v) for v in out)

    @property
    def mcs(self) -> int:
        return int(self.lib.orc_mcs(self._sim))

    @mcs.setter
    def mcs(self, v: int):
        self.lib.orc_set_mcs(self._sim, int(v))

    def set_seed(self, seed: int):
        self.lib.orc_set_seed(self._sim, int(seed))

    def rounds(self) -> int:
        return int(self.lib.orc_round_count(self._sim))

    def last_order(self) -> np.ndarray:
        out = np.zeros(256, np.uint32)
        n = self.lib.orc_last_order(self._sim, out.ctypes.data)
        return out[:n]

    def last_activations(self) -> np.ndarray:
        """Lottery: activated opportunities per mutable site in the last MCS."""
        out = np.zeros(int(np.prod(self.spec.dims)), np.uint32)
        n = self.lib.orc_last_activations(self._sim, out.ctypes.data)
        if n < 0:
            raise ValueError("activations exist for LotteryCPM only")
        return out[:n]

    def tracker_errors(self) -> int:
        return int(self.lib.orc_tracker_errors(self._sim))

    def total_energy(self) -> float:
        return float(self.lib.orc_total_energy(self._sim))

    def accepted_energy(self):
        """(sum of accepted delta_h over the last MCS, since the state was loaded), FP64."""
        out = np.zeros(2, np.float64)
        self.lib.orc_accepted_energy(self._sim, out.ctypes.data)
        return float(out[0]), float(out[1])

    def sample_transitions(self, seed_base: int, replicas: int, labels) -> np.ndarray:
        labels = np.ascontiguousarray(labels, dtype=np.int64)
        n = int(np.prod(self.spec.dims))
        counts = np.zeros(len(labels) ** n, np.uint64)
        rc = self.lib.orc_sample_transitions(self._sim, int(seed_base), int(replicas),
                                             labels.ctypes.data, len(labels), counts.ctypes.data)
        if rc != 0:
            raise RuntimeError("a destination owner is not among the catalog labels")
        return counts


def set_threads(threads: int) -> int:
    """Threads of the oracle's colour-pass loops (0: every processor); returns the count in use."""
    lib = load()
    lib.orc_set_threads(int(threads))
    return int(lib.orc_threads())


def coloring(spec) -> np.ndarray:
    lib = load()
    keep: list = []
    m = make_model(spec, keep)
    colors = np.zeros(int(np.prod(spec.dims)), np.uint8)
    maxdeg = C.c_int32()
    n = lib.orc_realized_coloring(C.byref(m), colors.ctypes.data, C.byref(maxdeg))
    if n < 0:
        raise RuntimeError("colouring exceeds 255 colours")
    return colors


def normal_box_muller(seed, stream, mcs, subround, operation, entity, invocation) -> np.float32:
    return np.float32(load().orc_normal_box_muller(int(seed), int(stream), int(mcs), int(subround), int(operation),
                                                   int(entity), int(invocation)))


def ou_transition(*args) -> np.float32:
    return np.float32(load().orc_ou_transition(*[float(a) for a in args]))


def rng_words(seed, stream, mcs=0, subround=0, operation=0, kind=0, entity=0, generation=0,
              invocation=0, draw=0):
    out = np.zeros(4, np.uint32)
    load().orc_rng_words(seed, stream, mcs, subround, operation, kind, entity, generation,
                         invocation, draw, out.ctypes.data)
    return tuple(int(v) for v in out)


def counter_key(seed, stream, mcs=0, subround=0, operation=0, kind=0, entity=0, generation=0,
                invocation=0, draw=0):
    out = np.zeros(3, np.uint64)
    load().orc_counter_key(seed, stream, mcs, subround, operation, kind, entity, generation,
                           invocation, draw, out.ctypes.data)
    return tuple(int(v) for v in out)


def philox(counter, key):
    c = np.asarray(counter, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    load().orc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return tuple(int(v) for v in out)


def claims(recipient, losing, gaining, priority, capacity: int):
    """The reference's three-phase cell-claim protocol on one colour (checkerboard.jl:273-369)
    over actionable attempts between cells; returns (selected, dispositions, max_priority,
    min_identity), the last two indexed by cell id."""
    lib = load()
    n = len(recipient)
    r = np.ascontiguousarray(recipient, np.uint32)
    l = np.ascontiguousarray(losing, np.uint32)
    g = np.ascontiguousarray(gaining, np.uint32)
    p = np.ascontiguousarray(priority, np.uint32)
    sel = np.zeros(n, np.uint8)
    disp = np.zeros(n, np.uint8)
    maxp = np.zeros(capacity + 1, np.uint32)  # indexed by cell id (entry 0 unused)
    mini = np.zeros(capacity + 1, np.uint32)
    lib.orc_claims(n, r.ctypes.data, l.ctypes.data, g.ctypes.data, p.ctypes.data, capacity,
                   sel.ctypes.data, disp.ctypes.data, maxp.ctypes.data, mini.ctypes.data)
    return sel, disp, maxp, mini
