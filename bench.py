#!/usr/bin/env python3
"""Headline benchmark: CheckerboardSweepCPM copy attempts per second on one MI355X.

Workload (one "step" = one MCS of the whole lattice): 4096 x 4096 periodic Int32 lattice,
163,840 cells of 10 x 10 seeded on a pitch-10 slot grid (rows filled in order), two cell
types, integer contact J, volume constraint (lambda 1, target 100), boundary constraint
(lambda 1, target 40), Moore contact relation, von Neumann proposal, T = 10, seed 1234.
Random numbers are the reference's Philox semantic addresses, so the run is synthetic but
fully deterministic. Timed region: `--steps` MCS after `--warmup` MCS, synchronized on both
sides. Attempts per MCS = activated_attempts of the report (one per mutable site).

Roofline: the sweep is HBM-bound stencil work. Its algorithmic traffic per MCS is one read of
the Int32 lattice per colour pass plus, per accepted copy, the lattice write and the volume and
boundary tracker updates: B_MCS = 4 * C * N + A * (4 + 16 + 32) bytes (C colour passes, N sites,
A accepted copies of the MCS). `roofline.achieved` = B_MCS x MCS/s of the timed run, so the
fraction covers every kernel of the MCS (propose, commit and the colour-order kernel), not one
of them. The per-kernel figures (propose: 4 * N bytes per launch over its average duration,
from HIP event pairs attached to every kernel's launch on the engine stream by
hipExtLaunchKernelGGL in cpm_step_timed) are under `extra.kernels`. `traffic` is the measured
HBM bytes of one MCS: C x (propose + commit) + the colour-order kernel, each from the
FETCH_SIZE and WRITE_SIZE hardware counters scaled by what the counters read for the
library's plain copy kernel of known size, taken from profiles/cb_4096_traffic.json when that
file is present.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "celly.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
METRIC = "copy attempts/sec + % HBM roofline, 1 GPU; batched vs plain; p99 job latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(passes: int, chunk_mcs: int):
    """Measured HBM bytes of one MCS (passes x (propose + commit), plus the plan and end launches
    of a chunk of chunk_mcs MCS shared over its MCS) from the committed counter summary, if
    present."""
    path = os.path.join(ROOT, "profiles", "cb_4096_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as handle:
        t = json.load(handle)
    try:
        return (passes * (t["propose"]["bytes_per_launch"] + t["commit"]["bytes_per_launch"])
                + (t["plan"]["bytes_per_launch"] + t["plan_end"]["bytes_per_launch"]) / max(chunk_mcs, 1))
    except KeyError:
        return None


def host_threads() -> int:
    """Host threads this process may use: OMP_NUM_THREADS when set, else the CPUs it may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "").strip()
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(sample_mcs: int, warmup_mcs: int):
    """The C restatement of the sweep (the oracle: colour passes on OpenMP threads, results
    independent of the thread count) on the same 4096^2 workload: `warmup_mcs` MCS from the
    initial state, then `sample_mcs` steady-state MCS on every host thread this process may use,
    then `sample_mcs` more on one thread. Returns (cpu_baseline, one-thread figure)."""
    from corepotts import build_spec, state_arrays, workloads
    from oracle.bridge import OracleSim, set_threads
    compiled, proposal, components, alg, seed = workloads.config2_large()
    spec = build_spec(compiled, proposal, components, alg, seed)
    sim = OracleSim(spec, state_arrays(compiled, spec))
    threads = set_threads(host_threads())
    sim.step(warmup_mcs)
    attempts_per_mcs = sim.report()[3]  # every mutable site is attempted once per MCS
    t0 = time.perf_counter()
    sim.step(sample_mcs)
    dt_all = time.perf_counter() - t0
    set_threads(1)
    t0 = time.perf_counter()
    sim.step(sample_mcs)
    dt_one = time.perf_counter() - t0
    set_threads(host_threads())
    base = {"value": attempts_per_mcs * sample_mcs / dt_all, "unit": "copy attempts/s",
            "cores": threads, "kind": "port",
            "sample": f"C restatement (oracle), {threads} OpenMP threads: MCS {warmup_mcs + 1}-"
                      f"{warmup_mcs + sample_mcs} of the 4096x4096 checkerboard workload after "
                      f"{warmup_mcs} warm-up MCS, {dt_all:.1f} s"}
    one = {"value": attempts_per_mcs * sample_mcs / dt_one, "unit": "copy attempts/s", "cores": 1,
           "sample": f"MCS {warmup_mcs + sample_mcs + 1}-{warmup_mcs + 2 * sample_mcs}, one thread, "
                     f"{dt_one:.1f} s"}
    return base, one


def state_digest(state: dict) -> str:
    import hashlib
    h = hashlib.blake2b(digest_size=16)
    for key in ("tags", "ids", "finite_volumes", "medium_volumes", "boundary_measures"):
        h.update(state[key].tobytes())
    return h.hexdigest()


def default_tile_shapes():
    """Tile-shape preferences back to the engine defaults (a tuned preference is process-wide
    and is tuned for one lattice)."""
    from corepotts import set_tile_shape
    for algorithm in (1, 2):
        for ndim in (2, 3):
            for halo in (1, 2):
                set_tile_shape(algorithm, ndim, halo, -1)


def batched_leg(replicas: int, mcs: int, graph_chunk: int, plain: bool = True, tune: bool = True,
                shape: int = -1):
    """Batched ensemble of `replicas` 2D replicas (sides log-uniform in [64, 1024], T, J and
    seed per replica) for `mcs` MCS in one ReplicaBatch, against the same replicas run one
    after another through the plain single-lattice path on the same GPU. shape >= 0 gives
    every size class that tile shape instead of tuning."""
    from corepotts import BatchShapes, ReplicaBatch, init_scientific, workloads
    from corepotts.spec import ALG_CHECKERBOARD
    log(f"batched: building {replicas} replicas")
    t0 = time.perf_counter()
    models = [workloads.ensemble_replica(r) for r in range(replicas)]
    sites = [int(m[0].tags.size) for m in models]
    log(f"batched: built in {time.perf_counter() - t0:.1f} s, {sum(sites) / 1e6:.1f} M sites")

    # tile shape, tuned on as many other draws of the same ensemble distribution (a smaller
    # sample weighs the fixed launch costs more and prefers larger tiles)
    t0 = time.perf_counter()
    shapes = BatchShapes()
    if shape >= 0:
        from corepotts.autotune import SIZE_CLASS_BOUNDS
        shapes.shapes = {(ALG_CHECKERBOARD, 2, 1, c): shape for c in range(len(SIZE_CLASS_BOUNDS) + 1)}
    elif tune:
        shapes.tune([workloads.ensemble_replica(replicas + i) for i in range(replicas)], per_class=replicas,
                    mcs=10, greedy=False)
    tune_s = time.perf_counter() - t0
    log(f"batched: shapes tuned in {tune_s:.1f} s: {shapes.describe()}")

    batch = ReplicaBatch(ALG_CHECKERBOARD, graph_chunk=graph_chunk, shapes=shapes)
    slots = [batch.add_model(*m, target_mcs=mcs) for m in models]
    batch.synchronize()
    log(f"batched: {mcs} MCS")
    t0 = time.perf_counter()
    batch.step(mcs)
    batch.synchronize()
    batched_s = time.perf_counter() - t0
    digests = [state_digest(batch.download(s)) for s in slots]
    device_bytes = batch.device_bytes()
    batch.close()

    log("plain: replicas one after another")
    default_tile_shapes()
    plain_s = 0.0
    same = 0
    for m, d in zip(models if plain else [], digests):
        integ = init_scientific(*m)
        integ.set_graph_chunk(graph_chunk)
        integ.synchronize()
        t0 = time.perf_counter()
        integ.step(mcs)
        integ.synchronize()
        plain_s += time.perf_counter() - t0
        same += state_digest(integ.download()) == d
        integ.close()
    attempts = sum(sites) * mcs
    return {"replicas": replicas, "mcs": mcs, "sites": sum(sites), "graph_chunk": graph_chunk,
            "batched_attempts_per_s": attempts / batched_s, "batched_s": batched_s,
            "plain_sequential_attempts_per_s": attempts / plain_s if plain else None,
            "plain_sequential_s": plain_s if plain else None,
            "speedup": plain_s / batched_s if plain else None,
            "replicas_identical_to_plain": same if plain else None,
            "device_bytes": device_bytes, "tile_shapes": shapes.describe(), "tune_s": tune_s}


def lottery_leg(steps: int, shape: int = -1):
    """4096^2 LotteryCPM on the headline workload (9 rounds per MCS: the Moore shell's conflict
    degree 8 + 1), tuned tile shape, then event-timed kernels."""
    from corepotts import init_scientific, set_tile_shape, tile_shapes, tune_tile_shape, workloads
    from corepotts.spec import ALG_LOTTERY
    log(f"lottery {steps} MCS")
    compiled, proposal, components, lalg, lseed = workloads.config2_large(algorithm="lottery")
    n = int(compiled.tags.size)
    if shape >= 0:
        set_tile_shape(ALG_LOTTERY, 2, 1, shape)
        ltuned = {"shape": shape, "tile": dict(tile_shapes(2, 1))[shape], "pinned": True}
    else:
        ltuned = tune_tile_shape(compiled, proposal, components, lalg, seed=lseed, mcs=2, warmup=1)
    lot = init_scientific(compiled, proposal, components, lalg, seed=lseed)
    # 6 warmup MCS (an even number of rounds), then graphs of the largest chunk of at most
    # 10 MCS that divides the timed steps, captured before the timed region
    lot.step(6)
    chunk = graph_chunk_for(steps, 10)
    lot.set_graph_chunk(chunk)
    prepared = lot.prepare_graph() if chunk else False
    lot.synchronize()
    t0 = time.perf_counter()
    lot.step(steps)
    lot.synchronize()
    dt = time.perf_counter() - t0
    lrep = lot.current_mcs_report()
    timing = lot.step_timed(2)
    trep = lot.current_mcs_report()
    propose_us = 1e3 * timing["propose_ms"] / max(timing["propose_launches"], 1)
    commit_us = 1e3 * timing["commit_ms"] / max(timing["commit_launches"], 1)
    mcs_bytes = 4.0 * trep.internal_rounds * n + trep.accepted_copies * (4 + 16 + 32)
    out = {"copy_attempts_per_s": lrep.activated_attempts * steps / dt,
           "mcs_per_s": steps / dt, "rounds": lrep.internal_rounds, "tile_shape": ltuned,
           "graph_chunk": chunk, "graphs_prepared": prepared,
           "propose_kernel_avg_us": propose_us, "commit_kernel_avg_us": commit_us,
           "algorithmic_bytes_per_mcs": mcs_bytes,
           "whole_mcs_algorithmic_GBps": mcs_bytes * steps / dt / 1e9,
           "last_report": lrep._asdict()}
    lot.close()
    return out


def cubes_leg(steps: int, graph_chunk: int, shape: int = -1):
    """256^3 CheckerboardSweepCPM: 32,768 cubes of 8^3 tiling the lattice, 3 types, Float32
    contact matrix U[2, 20], volume (1.5, 512) + 26-neighbour contact + boundary (0.5, 384),
    T = 20."""
    from corepotts import init_scientific, set_tile_shape, tile_shapes, tune_tile_shape, workloads
    from corepotts.spec import ALG_CHECKERBOARD
    log("3D: building 256^3 workload")
    model = workloads.cubes_3d(256)
    n = int(model[0].tags.size)
    if shape >= 0:
        set_tile_shape(ALG_CHECKERBOARD, 3, 1, shape)
        tuned = {"shape": shape, "tile": dict(tile_shapes(3, 1))[shape], "pinned": True}
    else:
        tuned = tune_tile_shape(*model, mcs=2, warmup=1)
    integ = init_scientific(*model)
    integ.set_graph_chunk(graph_chunk)
    integ.step(2 * graph_chunk)
    integ.synchronize()
    log(f"3D: timed {steps} MCS")
    t0 = time.perf_counter()
    integ.step(steps)
    integ.synchronize()
    dt = time.perf_counter() - t0
    rep = integ.current_mcs_report()
    timing = integ.step_timed(4)
    trep = integ.current_mcs_report()
    mcs_bytes = 4.0 * trep.internal_rounds * n + trep.accepted_copies * (4 + 16 + 32)
    out = {"lattice": [256, 256, 256], "cells": 32768, "colour_passes_per_mcs": rep.internal_rounds,
           "algorithmic_bytes_per_mcs": mcs_bytes,
           "whole_mcs_algorithmic_GBps": mcs_bytes * steps / dt / 1e9,
           "tile_shape": tuned,
           "copy_attempts_per_s": rep.activated_attempts * steps / dt, "mcs_per_s": steps / dt,
           "propose_kernel_avg_us": 1e3 * timing["propose_ms"] / max(timing["propose_launches"], 1),
           "commit_kernel_avg_us": 1e3 * timing["commit_ms"] / max(timing["commit_launches"], 1),
           "propose_algorithmic_GBps": 4.0 * n / (1e-3 * timing["propose_ms"] / max(timing["propose_launches"], 1)) / 1e9,
           "last_report": rep._asdict()}
    # FP64 energy bookkeeping over 20 more MCS (outside the timed region): the accepted copies'
    # delta_h summed in FP64 against the change of the global FP64 energy
    e0, acc = integ.total_energy(), 0.0
    for _ in range(20):
        integ.step(1)
        acc += integ.accepted_energy()
    e1 = integ.total_energy()
    out["energy_fp64"] = {"mcs": 20, "E_start": e0, "E_end": e1, "accepted_dh_sum": acc,
                          "drift": acc - (e1 - e0)}
    integ.close()
    return out


def tiled_leg(steps: int):
    """TiledCheckerboardCPM on the config-2 workload (4096 x 4096, 163,840 cells of 10 x 10,
    volume + Moore contact + boundary, T = 10) with the reference's default 4 x 4 tiles: one
    thread per tile runs the tile's 16 proposals of each colour in order."""
    from corepotts import init_scientific, workloads
    log("tiled: building the 4096^2 workload")
    model = workloads.config2_large(algorithm="tiled")
    integ = init_scientific(*model)
    integ.step(3)
    integ.synchronize()
    log(f"tiled: timed {steps} MCS")
    t0 = time.perf_counter()
    integ.step(steps)
    integ.synchronize()
    dt = time.perf_counter() - t0
    rep = integ.current_mcs_report()
    timing = integ.step_timed(4)
    batch_us = 1e3 * timing["propose_ms"] / max(timing["propose_launches"], 1)
    proposals_per_launch = rep.activated_attempts / max(rep.internal_rounds, 1)
    out = {"lattice": [4096, 4096], "cells": 163840, "tile": [4, 4],
           "internal_rounds": rep.internal_rounds,
           "copy_attempts_per_s": rep.activated_attempts * steps / dt, "mcs_per_s": steps / dt,
           "batch_kernel_avg_us": batch_us,
           "reconcile_kernel_avg_us": 1e3 * timing["commit_ms"] / max(timing["commit_launches"], 1),
           "proposals_per_batch_launch": proposals_per_launch,
           "proposals_per_s_in_batch_kernel": proposals_per_launch / (1e-6 * batch_us),
           "last_report": rep._asdict()}
    integ.close()
    return out


def serving_leg(jobs: int, chunk: int, budget_gb: float, graph_chunk: int = -1, rate: float = 200.0):
    """Served stream: the first `jobs` Poisson arrivals at `rate` jobs/s (80 % 2D side
    log-uniform [128, 2048], 20 % 3D side log-uniform [32, 128], 100-2000 MCS), admitted under
    an HBM budget and advanced by continuous batching (HIP graphs of `graph_chunk` MCS between
    membership changes; -1: one server step); requests are built before the clock starts,
    admission (compile, allocation, upload) and result download count."""
    from corepotts import BatchShapes, JobServer, latency_summary, make_request, workloads
    log(f"serving: building {jobs} requests")
    t0 = time.perf_counter()
    stream = workloads.serving_jobs(jobs, rate_per_s=rate)
    reqs = [make_request(j.index, j.arrival_s, *workloads.serving_model(j), mcs=j.mcs) for j in stream]
    log(f"serving: built in {time.perf_counter() - t0:.1f} s")
    # tile shape per size class, tuned before serving on another draw of the stream
    t0 = time.perf_counter()
    shapes = BatchShapes()
    sample = workloads.serving_jobs(64, seed=workloads.SERVING_SEED + 1)
    shapes.tune([workloads.serving_model(j) for j in sample], per_class=6, mcs=10)
    tune_s = time.perf_counter() - t0
    log(f"serving: shapes tuned in {tune_s:.1f} s: {shapes.describe()}")
    server = JobServer(int(budget_gb * 1e9), chunk_mcs=chunk, shapes=shapes,
                       graph_chunk=None if graph_chunk < 0 else graph_chunk)
    t0 = time.perf_counter()
    done = server.serve(reqs)
    wall = time.perf_counter() - t0
    summary = latency_summary(done)
    summary.update({"wall_s": wall, "chunk_mcs": chunk, "budget_bytes": int(budget_gb * 1e9),
                    "peak_jobs_resident": server.peak_jobs, "peak_bytes": server.peak_used,
                    "arrival_rate_per_s": rate, "jobs_3d": sum(j.ndim == 3 for j in stream),
                    "graph_chunk": server.graph_chunk, "tile_shapes": shapes.describe(),
                    "tune_s": tune_s})
    server.batch.close()
    return summary


def graph_chunk_for(steps: int, limit: int) -> int:
    """Largest divisor of `steps` that is at most `limit` (0: no graphs)."""
    if limit <= 0 or steps <= 0:
        return 0
    return max(c for c in range(1, min(steps, limit) + 1) if steps % c == 0)


def headline_leg(args):
    """The 4096^2 CheckerboardSweepCPM run: tuned tile shape, `--warmup` MCS, then `--steps`
    timed MCS with HIP graphs over `--graph-chunk` MCS chunks; then event-timed kernels."""
    from corepotts import copy_probe, init_scientific, set_tile_shape, tile_shapes, tune_tile_shape, workloads
    from corepotts.spec import ALG_CHECKERBOARD

    # plain device copy of known size: achievable-bandwidth figure and the calibration
    # kernel for the traffic counters of a profiled run
    copy_bytes = args.copy_probe_mib << 20
    copy_ms = copy_probe(copy_bytes, 10)

    log("building 4096x4096 workload")
    compiled, proposal, components, alg, seed = workloads.config2_large()
    n_sites = int(compiled.tags.size)
    if args.tile_shape >= 0:
        set_tile_shape(ALG_CHECKERBOARD, 2, 1, args.tile_shape)
        tuned = {"shape": args.tile_shape, "tile": dict(tile_shapes(2, 1))[args.tile_shape], "pinned": True}
    else:
        log("tuning the tile shape")
        tuned = tune_tile_shape(compiled, proposal, components, alg, seed=seed)
    log(f"tile shape {tuned['shape']} {tuned['tile']}; compiling schedule and uploading")
    integ = init_scientific(compiled, proposal, components, alg, seed=seed)
    integ.set_graph_chunk(args.graph_chunk)
    log(f"warmup {args.warmup} MCS")
    integ.step(args.warmup)
    integ.synchronize()
    # graphs of the largest chunk (at most --graph-chunk MCS) that divides the timed steps,
    # captured before the timed region: every timed MCS replays a graph, however few there are
    chunk = graph_chunk_for(args.steps, args.graph_chunk)
    integ.set_graph_chunk(chunk)
    graphs_ready = integ.prepare_graph() if chunk else False
    integ.synchronize()

    log(f"timed {args.steps} MCS (graphs of {chunk} MCS, prepared: {graphs_ready})")
    t0 = time.perf_counter()
    integ.step(args.steps)
    integ.synchronize()
    elapsed = time.perf_counter() - t0
    rep = integ.current_mcs_report()
    attempts_per_mcs = rep.activated_attempts
    ms_per_step = 1000.0 * elapsed / args.steps
    value = attempts_per_mcs * args.steps / elapsed

    # host <-> device transfers at the boundary (outside the timed region): one state upload
    # and one download, to state the PCIe-inclusive rate of the timed run
    t0 = time.perf_counter()
    integ.upload(compiled)
    integ.synchronize()
    upload_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    integ.download()
    download_s = time.perf_counter() - t0

    log(f"kernel timing over {args.timing_mcs} MCS")
    timing = integ.step_timed(args.timing_mcs)
    propose_avg_ms = timing["propose_ms"] / max(timing["propose_launches"], 1)
    commit_avg_ms = timing["commit_ms"] / max(timing["commit_launches"], 1)
    # the MCS-order kernels: one plan launch and one end launch per chunk of up to 64 MCS
    order_ms_per_mcs = timing["begin_ms"] / max(timing["mcs"], 1)
    rep_t = integ.current_mcs_report()
    passes = rep_t.internal_rounds
    # algorithmic bytes of one MCS: a lattice read per colour pass + the accepted copies'
    # lattice write and tracker updates (4 + 2 x 4 volume RMW + 2 x 8 boundary RMW, read and
    # write: 4 + 16 + 32 bytes)
    mcs_bytes = 4.0 * passes * n_sites + rep_t.accepted_copies * (4 + 16 + 32)
    mcs_per_s = args.steps / elapsed
    achieved = mcs_bytes * mcs_per_s / 1e9
    traffic = traffic_from_profiles(passes, chunk if chunk > 0 else 1)
    propose_gbps = 4.0 * n_sites / (propose_avg_ms * 1e-3) / 1e9

    extra = {
        "mcs_per_s": mcs_per_s,
        "graph_chunk": chunk,
        "graphs_prepared": graphs_ready,
        "tile_shape": tuned,
        "pcie": {"upload_ms": upload_s * 1e3, "download_ms": download_s * 1e3,
                 "pcie_inclusive_attempts_per_s":
                     attempts_per_mcs * args.steps / (elapsed + upload_s + download_s)},
        "copy_probe": {"bytes": copy_bytes, "avg_ms": copy_ms,
                       "GBps": 2.0 * copy_bytes / (copy_ms * 1e-3) / 1e9},
        "colour_passes_per_mcs": passes,
        "algorithmic_bytes_per_mcs": mcs_bytes,
        "kernels": {
            "propose_avg_us": propose_avg_ms * 1e3,
            "propose_algorithmic_GBps": propose_gbps,
            "propose_frac_of_peak": propose_gbps / HBM_PEAK_GBS,
            "commit_avg_us": commit_avg_ms * 1e3,
            "mcs_order_us_per_mcs": order_ms_per_mcs * 1e3,
            "event_wall_ms_per_mcs": timing["wall_ms"] / max(timing["mcs"], 1),
        },
        "last_report": rep._asdict(),
    }
    del integ
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "copy attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+fp32",
        "data": "synthetic (deterministic seeded layout, Philox semantic RNG)",
        "config": {"workload": "CheckerboardSweepCPM 4096x4096 periodic, 163840 cells 10x10, "
                               "volume+Moore contact+boundary, T=10, seed 1234",
                   "lattice": [4096, 4096], "cells": 163840, "algorithm": "CheckerboardSweepCPM"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "scope": "whole MCS: B_MCS x MCS/s, B_MCS = 4*C*N + A*52 bytes"},
    }
    return line, extra


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--graph-chunk", type=int, default=50)
    ap.add_argument("--timing-mcs", type=int, default=40)
    ap.add_argument("--lottery-steps", type=int, default=50)
    ap.add_argument("--cpu-sample-mcs", type=int, default=20)
    ap.add_argument("--cpu-warmup-mcs", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-lottery", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=256)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-3d", action="store_true")
    ap.add_argument("--cubes-steps", type=int, default=40)
    ap.add_argument("--batch-replicas", type=int, default=512)
    ap.add_argument("--batch-mcs", type=int, default=500)
    ap.add_argument("--batch-no-plain", action="store_true",
                    help="skip the sequential plain runs of the batched leg (profiling the batch alone)")
    ap.add_argument("--batch-no-tune", action="store_true",
                    help="default tile shapes for the batched leg instead of tuning them")
    ap.add_argument("--batch-tile-shape", type=int, default=-1,
                    help="one tile shape for every size class of the batched leg instead of tuning")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-tiled", action="store_true")
    ap.add_argument("--no-headline", action="store_true",
                    help="skip the 4096^2 checkerboard run (profiling the other legs alone)")
    ap.add_argument("--tiled-steps", type=int, default=20)
    ap.add_argument("--serve-jobs", type=int, default=200)
    ap.add_argument("--serve-chunk", type=int, default=20)
    ap.add_argument("--serve-graph-chunk", type=int, default=-1,
                    help="MCS per HIP graph in the served batch (-1: one server step, 0: no graphs)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="job arrivals per second")
    ap.add_argument("--tile-shape", type=int, default=-1,
                    help="2D checkerboard tile shape of the headline run (default: autotuned)")
    ap.add_argument("--lottery-tile-shape", type=int, default=-1,
                    help="2D lottery tile shape (default: autotuned)")
    ap.add_argument("--cubes-tile-shape", type=int, default=-1,
                    help="3D checkerboard tile shape (default: autotuned)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this engine runs one lattice (or one batch of "
              "replicas) per GPU with no sharding or collectives; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from corepotts import init_scientific, tune_tile_shape, workloads

    line = {"metric": METRIC, "n_gpus": 1, "vs_baseline": None}
    extra = {}
    if not args.no_headline:
        line, extra = headline_leg(args)

    if not args.no_lottery:
        extra["lottery"] = lottery_leg(args.lottery_steps, args.lottery_tile_shape)

    if not args.no_3d:
        extra["cubes_3d"] = cubes_leg(args.cubes_steps, 10, args.cubes_tile_shape)
    if not args.no_tiled:
        extra["tiled"] = tiled_leg(args.tiled_steps)
    if not args.no_batch:
        extra["batched"] = batched_leg(args.batch_replicas, args.batch_mcs, args.graph_chunk,
                                       not args.batch_no_plain, not args.batch_no_tune, args.batch_tile_shape)
    if not args.no_serve:
        extra["serving"] = serving_leg(args.serve_jobs, args.serve_chunk, 200.0, args.serve_graph_chunk,
                                       args.serve_rate)

    baseline = None
    if not args.no_cpu_baseline:
        log("cpu baseline (C restatement, every host thread, then one thread)")
        baseline, one_thread = cpu_baseline(args.cpu_sample_mcs, args.cpu_warmup_mcs)
        extra["cpu_baseline_one_thread"] = one_thread
        if line.get("value"):
            extra["gpu_over_cpu"] = {"all_threads": line["value"] / baseline["value"],
                                     "one_thread": line["value"] / one_thread["value"]}

    line["cpu_baseline"] = baseline
    line["extra"] = extra
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
