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
