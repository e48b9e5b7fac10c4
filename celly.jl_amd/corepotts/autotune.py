"""Tile-shape autotuner for the propose kernels.

Each propose launch processes one tile of the lattice per workgroup; the tile shape sets
the halo overhead, the sites per thread and the LDS footprint (hence occupancy), and the
best shape differs by algorithm (checkerboard passes touch one colour of a tile, lottery
rounds draw a ticket for every tile position) and dimension. The tuner times every
compiled shape that matches a model's dimension and halo on the model itself and keeps
the fastest as the process-wide preference for replicas created afterwards. The shape
never changes results, only speed.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Tuple

from .native import EngineError, init_scientific, load_library
from .spec import build_spec


def tile_shapes(ndim: int, halo: int) -> List[Tuple[int, Tuple[int, int, int]]]:
    """(shape id, (TX, TY, TZ)) of every compiled tile shape for this dimension and halo."""
    lib = load_library()
    out = []
    info = (C.c_int32 * 5)()
    for shape in range(16):
        if lib.cpm_tile_shape_info(shape, info) == 0 and info[0] == ndim and info[1] == halo:
            out.append((shape, (info[2], info[3], info[4])))
    return out


def set_tile_shape(algorithm: int, ndim: int, halo: int, shape: int):
    """Preferred shape for replicas of (algorithm, ndim, halo) created afterwards (-1: default)."""
    rc = load_library().cpm_set_tile_shape(int(algorithm), int(ndim), int(halo), int(shape))
    if rc != 0:
        raise EngineError(rc, "tile shape does not match the dimension and halo")


def set_row_kernels(enable: bool):
    """Row propose kernels (2D periodic von Neumann checkerboards with a 2 x 2 colour pattern)
    on or off for replicas created afterwards; on by default. Results do not depend on it."""
    load_library().cpm_set_row_kernels(1 if enable else 0)


def set_tiled_staging(enable: bool):
    """The staged TiledCheckerboardCPM kernel (tile box and halo in LDS) on or off for models
    created afterwards; on by default. Results do not depend on it."""
    load_library().cpm_set_tiled_staging(1 if enable else 0)


def model_halo(spec) -> int:
    """Largest |offset| component over the relations the kernels stage (at least 1)."""
    h = 1
    for rel in (spec.proposal, spec.contact, spec.surface):
        if rel is not None:
            for o in rel.offsets:
                h = max(h, max(abs(int(c)) for c in o))
    return h


def tune_tile_shape(state, proposal_relation, components, algorithm, seed: int = 0,
                    mcs: int = 4, warmup: int = 2) -> Dict:
    """Times each candidate shape on this model (event-timed MCS, after warmup MCS) and sets
    the fastest as the preference. Returns {"shape", "tile", "ms_per_mcs": {shape: ms},
    "row_kernel": {shape: whether the row propose kernel ran}}."""
    spec = build_spec(state, proposal_relation, components, algorithm, seed)
    ndim, halo = spec.ndim, model_halo(spec)
    timings = {}
    tiles = {}
    rows = {}
    for shape, tile in tile_shapes(ndim, halo):
        set_tile_shape(spec.algorithm, ndim, halo, shape)
        integ = init_scientific(state, proposal_relation, components, algorithm, seed=seed)
        rows[shape] = bool(integ.kernel_config() & 32)
        integ.step(warmup)
        t = integ.step_timed(mcs)
        integ.close()
        timings[shape] = t["wall_ms"] / max(t["mcs"], 1)
        tiles[shape] = tile
    best = min(timings, key=timings.get)
    set_tile_shape(spec.algorithm, ndim, halo, best)
    return {"shape": best, "tile": list(tiles[best]), "ms_per_mcs": timings, "row_kernel": rows}


# Size classes of batched replicas, by the longest lattice side: one tile shape is tuned per
# (algorithm, dimension, halo, class). Replicas of different shapes form different launch groups.
SIZE_CLASS_BOUNDS = (128, 256, 512)


def size_class(spec) -> int:
    side = max(int(d) for d in spec.dims[:spec.ndim])
    return sum(side >= b for b in SIZE_CLASS_BOUNDS)


class BatchShapes:
    """Tile shape per (algorithm, dimension, halo, size class) for batched replicas.

    Replicas of different tile shapes run in different launch groups, so a class's best shape
    depends on the others (one group for everything saves launches; a shape fitted to a class
    saves halo work). tune() therefore times the whole sample batch (HIP graphs of `mcs` MCS,
    timed after a warm-up step of the same length). It first gives every class of a
    (algorithm, dimension, halo) family the same shape, tries every compiled shape that way
    and starts from the fastest; then it goes class by class, from the largest lattices down,
    tries every shape for that class with the other classes fixed, and keeps the fastest.
    apply(spec) sets the process-wide preference for the next replica of that class;
    ReplicaBatch.reserve calls it. Shapes change speed only, never results."""

    def __init__(self):
        self.shapes: Dict[tuple, int] = {}
        self.timings: Dict[tuple, Dict[int, float]] = {}
        self.uniform_timings: Dict[tuple, Dict[int, float]] = {}

    @staticmethod
    def key(spec) -> tuple:
        return (spec.algorithm, spec.ndim, model_halo(spec), size_class(spec))

    def apply(self, spec):
        k = self.key(spec)
        set_tile_shape(spec.algorithm, spec.ndim, k[2], self.shapes.get(k, -1))

    def _time_sample(self, sample, mcs: int) -> float:
        import time
        from .batch import ReplicaBatch
        batch = ReplicaBatch(sample[0][1].algorithm, graph_chunk=mcs, shapes=self)
        try:
            for m, spec in sample:
                batch.add(m[0], spec, target_mcs=2 * mcs)
            batch.step(mcs)  # warm-up: first launches and graph capture
            batch.synchronize()
            t0 = time.perf_counter()
            batch.step(mcs)
            batch.synchronize()
            return time.perf_counter() - t0
        finally:
            batch.close()

    def tune(self, models, per_class: int = 16, mcs: int = 10, greedy: bool = True) -> Dict[tuple, int]:
        """models: init_scientific argument tuples of one algorithm; the first per_class
        replicas of each class form the timed sample. greedy=False keeps the best shared shape
        (no class-by-class step): for a sample as large as the batch it serves, whose fixed
        launch costs weigh as they will there."""
        classes: Dict[tuple, list] = {}
        for m in models:
            spec = build_spec(*m)
            members = classes.setdefault(self.key(spec), [])
            if len(members) < per_class:
                members.append((m, spec))
        sample = [x for members in classes.values() for x in members]
        if not sample:
            return self.shapes
        for fam in sorted({k[:3] for k in classes}):
            members = [k for k in classes if k[:3] == fam]
            times = {}
            for shape, _tile in tile_shapes(fam[1], fam[2]):
                for k in members:
                    self.shapes[k] = shape
                times[shape] = self._time_sample(sample, mcs)
            for k in members:
                self.shapes[k] = min(times, key=times.get)
            self.uniform_timings[fam] = times
        for k in sorted(classes, key=lambda k: -k[3]) if greedy else ():
            times = {}
            for shape, _tile in tile_shapes(k[1], k[2]):
                self.shapes[k] = shape
                times[shape] = self._time_sample(sample, mcs)
            self.shapes[k] = min(times, key=times.get)
            self.timings[k] = times
        for k in classes:  # the process-wide preferences back to the defaults
            set_tile_shape(k[0], k[1], k[2], -1)
        return self.shapes

    def describe(self) -> Dict[str, dict]:
        """{"alg/ndim/halo/class": {"shape", "tile", "ms": {shape: ms}}} for reports."""
        out = {}
        for k, shape in self.shapes.items():
            tiles = dict(tile_shapes(k[1], k[2]))
            out["%d/%dd/h%d/c%d" % k] = {
                "shape": shape, "tile": list(tiles[shape]),
                "ms": {int(s): round(t * 1e3, 3) for s, t in self.timings.get(k, {}).items()}}
        for fam, times in self.uniform_timings.items():
            out["%d/%dd/h%d/every class" % fam] = {
                "ms": {int(s): round(t * 1e3, 3) for s, t in times.items()}}
        return out
