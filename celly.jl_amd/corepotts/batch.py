"""Many independent replicas on one GPU, advanced together (cpm_batch_* of the C-ABI).

A batch plays the role of the reference's EnsembleProblem trajectories
(lib/CorePotts/src/sciml/interface.jl:1059-1084): every replica is an independent
ScientificPottsIntegrator with its own lattice size, temperature, contact matrix and seed,
and its trajectory equals the one a single integrator produces from the same state and seed.
On the device the replicas of one tile shape share each colour-pass launch, so a step of
the batch costs one begin launch plus three launches (propose, claim check, apply) per
colour pass, whatever the replica count. Replicas may be added and removed between steps (continuous batching).
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np

from .model import CompiledScientificState
from .native import (EngineError, ScientificMCSReport, _ptr, load_library, make_desc, mechanics_arrays,
                     mechanics_columns, snapshot, state_buffers, state_pointers)
from .spec import ALG_CHECKERBOARD, ALG_LOTTERY, ALG_TILED, ModelSpec, build_spec, state_arrays


def replica_bytes(spec: ModelSpec) -> int:
    """Device bytes one replica of this model holds (cpm_replica_bytes), for admission."""
    lib = load_library()
    keep: list = []
    desc = make_desc(spec, keep)
    out = C.c_uint64()
    rc = lib.cpm_replica_bytes(C.byref(desc), C.byref(out))
    if rc != 0:
        raise EngineError(rc, "cpm_replica_bytes failed")
    return int(out.value)


class ReplicaBatch:
    """Device-resident replicas of one algorithm (CheckerboardSweepCPM, LotteryCPM or
    TiledCheckerboardCPM)."""

    def __init__(self, algorithm: int = ALG_CHECKERBOARD, hbm_budget_bytes: int = 0,
                 graph_chunk: int = 0, shapes=None):
        if algorithm not in (ALG_CHECKERBOARD, ALG_LOTTERY, ALG_TILED):
            raise NotImplementedError("the MI355X engine batches CheckerboardSweepCPM, LotteryCPM and "
                                      "TiledCheckerboardCPM")
        self.lib = load_library()
        self.algorithm = algorithm
        handle = C.c_void_p()
        rc = self.lib.cpm_batch_create(int(algorithm), int(hbm_budget_bytes), C.byref(handle))
        if rc != 0:
            raise EngineError(rc, self.lib.cpm_batch_last_error(None).decode())
        self._b = handle
        self.shapes = shapes  # optional autotune.BatchShapes: tile shape per replica size class
        self.specs: Dict[int, ModelSpec] = {}
        self.compiled: Dict[int, CompiledScientificState] = {}
        if graph_chunk:
            self.set_graph_chunk(graph_chunk)

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "_b", None):
            self.lib.cpm_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            raise EngineError(rc, self.lib.cpm_batch_last_error(self._b).decode())

    # -- membership --------------------------------------------------------
    def reserve(self, spec: ModelSpec) -> int:
        """Compile the schedule of a replica and allocate its device state (cpm_batch_add);
        returns its slot. No host synchronization: the host work overlaps queued GPU work.
        Raises EngineError with code CPM_E_NOMEM when the HBM budget would be exceeded. The
        replica stays idle until load() gives it a state and a target."""
        if spec.algorithm != self.algorithm:
            raise ValueError("replica algorithm differs from the batch")
        if self.shapes is not None:
            self.shapes.apply(spec)
        keep: list = []
        desc = make_desc(spec, keep)
        slot = C.c_int32()
        self._check(self.lib.cpm_batch_add(self._b, C.byref(desc), C.byref(slot)))
        self.specs[int(slot.value)] = spec
        return int(slot.value)

    def load(self, slot: int, compiled: CompiledScientificState, target_mcs: Optional[int] = None,
             initialize_mechanics: bool = True):
        """Upload the state of a reserved replica and set its target MCS. Mechanical states are
        initialized as init_scientific does unless initialize_mechanics is False."""
        spec = self.specs[slot]
        try:
            a = state_arrays(compiled, spec)
            self._check(self.lib.cpm_batch_upload_state(self._b, int(slot), *state_pointers(a)))
            if spec.mechanics:
                qp, qt = mechanics_arrays(compiled, spec)
                self._check(self.lib.cpm_batch_upload_mechanics(self._b, int(slot), _ptr(qp), _ptr(qt)))
                if initialize_mechanics:
                    self._check(self.lib.cpm_batch_initialize_mechanics(self._b, int(slot)))
            if target_mcs is not None:
                self.set_target(slot, target_mcs)
        except Exception:
            self.remove(slot)
            raise
        self.compiled[slot] = compiled

    def add(self, compiled: CompiledScientificState, spec: ModelSpec,
            target_mcs: Optional[int] = None) -> int:
        """reserve() then load(): compile, allocate and upload one replica; returns its slot."""
        slot = self.reserve(spec)
        self.load(slot, compiled, target_mcs)
        return slot

    def add_model(self, state, proposal_relation, components, algorithm, seed: int = 0,
                  target_mcs: Optional[int] = None) -> int:
        """init_scientific's arguments (checkerboard.jl:197-219) as one more replica."""
        spec = build_spec(state, proposal_relation, components, algorithm, seed)
        return self.add(state, spec, target_mcs)

    def checkpoint(self, slot: int):
        """CanonicalCheckpoint of one replica at its last completed MCS (synchronizes): the
        state, seed and MCS counter from which resume() continues it exactly."""
        from .checkpoint import capture_batch_checkpoint
        return capture_batch_checkpoint(self, slot)

    def resume(self, slot: int, cp, prototype_compiled: CompiledScientificState,
               target_mcs: Optional[int] = None):
        """Load a reserved replica from a checkpoint of its model (persistence/scientific.jl:
        464-468): ownership, trackers rebuilt from it, mechanical states as captured, the
        checkpoint's MCS counter and the reserved spec's seed (the checkpoint's own)."""
        from .checkpoint import CheckpointCompatibilityError, checkpoint_compiled_state, \
            model_fingerprint, topology_fingerprint
        spec = self.specs[slot]
        if cp.topology_fingerprint != topology_fingerprint(spec) or cp.model_fingerprint != model_fingerprint(spec):
            raise CheckpointCompatibilityError("model", "replica model fingerprint", "checkpoint model fingerprint")
        if int(cp.seed) != int(spec.seed):
            raise CheckpointCompatibilityError("seed", str(spec.seed), str(cp.seed))
        self.load(slot, checkpoint_compiled_state(cp, prototype_compiled), initialize_mechanics=False)
        self.set_mcs(slot, int(cp.mcs))
        if target_mcs is not None:
            self.set_target(slot, target_mcs)

    @staticmethod
    def snapshot_bytes(spec: ModelSpec) -> int:
        """Device bytes of a snapshot of a replica of this model: the owner words, the 64-byte
        cell records (capacity + 1) and the medium-volume counters (cpm_snapshot_info)."""
        return 4 * int(np.prod(spec.dims)) + 64 * (int(spec.capacity) + 1) + 4 * 16

    def snapshot(self, slot: int):
        """Device-resident snapshot of one replica at its last completed MCS (synchronizes; no
        host transfer): restore_snapshot() continues it exactly in a reserved slot."""
        from .native import DeviceSnapshot
        h = C.c_void_p()
        self._check(self.lib.cpm_batch_snapshot_capture(self._b, int(slot), C.byref(h)))
        return DeviceSnapshot(self.lib, h, self.specs[slot])

    def restore_snapshot(self, slot: int, snap, prototype_compiled: Optional[CompiledScientificState] = None,
                         target_mcs: Optional[int] = None):
        """Load a reserved replica of the snapshot's model from a device snapshot (state,
        seed and MCS counter, copied on the device) and set its target. prototype_compiled
        (the model's compiled state) serves logical_state() afterwards."""
        spec = self.specs[slot]
        if int(snap.seed) != int(spec.seed):
            raise ValueError(f"snapshot seed {snap.seed} differs from the replica's {spec.seed}")
        self._check(self.lib.cpm_batch_snapshot_restore(self._b, int(slot), snap.handle))
        if prototype_compiled is not None:
            self.compiled[slot] = prototype_compiled
        if target_mcs is not None:
            self.set_target(slot, target_mcs)

    def remove(self, slot: int):
        self._check(self.lib.cpm_batch_remove(self._b, int(slot)))
        self.specs.pop(slot, None)
        self.compiled.pop(slot, None)

    @property
    def slots(self):
        return sorted(self.specs)

    # -- stepping ----------------------------------------------------------
    def set_target(self, slot: int, target_mcs: int):
        """The replica advances (in later steps) until its MCS counter reaches target_mcs."""
        self._check(self.lib.cpm_batch_set_target(self._b, int(slot), int(target_mcs)))

    def set_mcs(self, slot: int, mcs: int):
        """Set a replica's MCS counter (checkpoint restore); it idles until set_target."""
        self._check(self.lib.cpm_batch_set_mcs(self._b, int(slot), int(mcs)))

    def step(self, n: int = 1):
        """Enqueue n MCS; each replica below its target performs min(n, remaining) of them."""
        if n < 0:
            raise ValueError("step count must be non-negative")
        self._check(self.lib.cpm_batch_step(self._b, int(n)))

    def set_graph_chunk(self, chunk: int):
        self._check(self.lib.cpm_batch_set_graph_chunk(self._b, int(chunk)))

    def synchronize(self):
        self._check(self.lib.cpm_batch_sync(self._b))

    def progress(self) -> Dict[int, int]:
        """Current MCS of every replica (synchronizes)."""
        cap = max(int(self.lib.cpm_batch_capacity(self._b)), 1)
        out = np.zeros(cap, dtype=np.uint64)
        self._check(self.lib.cpm_batch_progress(self._b, out.ctypes.data, cap))
        return {s: int(out[s]) for s in self.specs}

    # -- observation -------------------------------------------------------
    def download(self, slot: int) -> dict:
        spec = self.specs[slot]
        out = state_buffers(spec)
        self._check(self.lib.cpm_batch_download_state(self._b, int(slot), *state_pointers(out)))
        out["medium_volumes"] = out["medium_volumes"][:len(spec.medium_ids)]
        if spec.mechanics:
            qp, qt = (None if key is None else np.zeros(spec.capacity, np.float32)
                      for key in mechanics_columns(spec))
            self._check(self.lib.cpm_batch_download_mechanics(self._b, int(slot), _ptr(qp), _ptr(qt)))
            for key, arr in zip(mechanics_columns(spec), (qp, qt)):
                if key is not None:
                    out[key] = arr
        return out

    def logical_state(self, slot: int) -> CompiledScientificState:
        return snapshot(self.compiled[slot], self.download(slot))

    def report(self, slot: int) -> ScientificMCSReport:
        out = (C.c_uint64 * 12)()
        self._check(self.lib.cpm_batch_report(self._b, int(slot), out))
        return ScientificMCSReport(*[int(v) for v in out])

    def device_bytes(self) -> int:
        out = C.c_uint64()
        self._check(self.lib.cpm_batch_bytes(self._b, C.byref(out)))
        return int(out.value)
