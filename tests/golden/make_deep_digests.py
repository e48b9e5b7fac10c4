"""Writes tests/golden/deep_digests.json: oracle state digests and reports of the configured
workloads at depth, the fixture the GPU runs of tests/test_gpu_deep.py are compared with.

Parts (each a list of {mcs, digest, report}):
  config2_checkerboard  4096^2, 163,840 cells, volume + Moore contact + boundary, T = 10,
                        seed 1234: MCS 10, 100, 1000
  config2_lottery       the same model under LotteryCPM: MCS 10, 100
  config3_checkerboard  256^3, 32,768 cubes, Float32 J, 26-neighbour contact: MCS 10, 200
  config4_ensemble      the 512-replica batched ensemble: every replica at MCS 50
  config5_jobs          eight served jobs at configured sizes (six 2D of side >= 1024, two 3D of
                        side >= 96): each at MCS 20

Run from the repository root: python tests/golden/make_deep_digests.py [--parts a,b] [--workers 8]
The three large parts run one after another with the oracle's colour-pass loops on every
processor (the oracle's results do not depend on its thread count); ensemble replicas and jobs
run single-threaded in parallel processes. The file is rewritten after every part, and parts
already in it are kept unless named again. The full set takes about 20 minutes on 8 cores.
"""
import argparse
import json
import multiprocessing
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "celly.jl_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

OUT = os.path.join(HERE, "deep_digests.json")


def _run(model, checkpoints):
    from helpers import oracle_for, state_digest
    sim, _ = oracle_for(*model)
    done = 0
    out = []
    for target in checkpoints:
        sim.step(target - done)
        done = target
        out.append({"mcs": target, "digest": state_digest(sim.store()), "report": list(sim.report())})
    return out


def part(name):
    from corepotts import workloads
    t0 = time.time()
    if name == "config2_checkerboard":
        res = _run(workloads.config2_large(), [10, 100, 1000])
    elif name == "config2_lottery":
        res = _run(workloads.config2_large(algorithm="lottery"), [10, 100])
    elif name == "config3_checkerboard":
        res = _run(workloads.cubes_3d(256), [10, 200])
    elif name.startswith("config4_replica_"):
        r = int(name.rsplit("_", 1)[1])
        res = _run(workloads.ensemble_replica(r), [50])
    elif name.startswith("config5_job_"):
        from helpers import deep_serving_jobs
        k = int(name.rsplit("_", 1)[1])
        job = deep_serving_jobs()[k]
        res = _run(workloads.serving_model(job), [job.mcs])
    else:
        raise ValueError(name)
    print(f"{name}: {time.time() - t0:.1f} s", flush=True)
    return name, res


LARGE = ("config2_checkerboard", "config2_lottery", "config3_checkerboard")


def _single_threaded(name):
    from oracle.bridge import set_threads
    set_threads(1)
    return part(name)


def _store(data, name, res):
    if name.startswith("config4_replica_"):
        data.setdefault("config4_ensemble", {})[name.rsplit("_", 1)[1]] = res
    elif name.startswith("config5_job_"):
        data.setdefault("config5_jobs", {})[name.rsplit("_", 1)[1]] = res
    else:
        data[name] = res


def _write(data):
    tmp = OUT + ".tmp"
    with open(tmp, "w") as handle:
        json.dump(data, handle, indent=0, sort_keys=True)
    os.replace(tmp, OUT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--parts", default="config2_checkerboard,config2_lottery,config3_checkerboard,config4,config5")
    ap.add_argument("--workers", type=int, default=8)
    args = ap.parse_args()
    names = []
    for p in args.parts.split(","):
        if p == "config4":
            names += [f"config4_replica_{r}" for r in range(512)]
        elif p == "config5":
            names += [f"config5_job_{k}" for k in range(8)]
        else:
            names.append(p)
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as handle:
            data = json.load(handle)
    from oracle.bridge import set_threads
    set_threads(0)
    for name in [n for n in names if n in LARGE]:
        _store(data, *part(name))
        _write(data)
    small = [n for n in names if n not in LARGE]
    if small:
        # spawned, not forked: the parent has run OpenMP regions
        ctx = multiprocessing.get_context("spawn")
        with ProcessPoolExecutor(max_workers=args.workers, mp_context=ctx) as pool:
            for name, res in pool.map(_single_threaded, small):
                _store(data, name, res)
        _write(data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
