"""Summaries of rocprofv3 output for the engine's kernels.

  python -m corepotts.rocprof_summary stats <kernel_stats.csv> [--top N]
      markdown table of `rocprofv3 --kernel-trace --stats` (calls, average/min/max duration, share)
  python -m corepotts.rocprof_summary counters <pmc_counter_collection.csv>... [--match SUBSTR]
      markdown table of SQ counters per wave: each counter summed over a kernel's dispatches and
      divided by the summed SQ_WAVES (one --pmc pass per file; every pass must hold SQ_WAVES)
  python -m corepotts.rocprof_summary ratios <pmc_counter_collection.csv> [--match SUBSTR]
      per-kernel sums per dispatch of every counter of one pass (atomic, TCP and TCC counters)
  python -m corepotts.rocprof_summary traffic --fetch <csv> --write <csv> --copy-bytes B
          --kernel NAME=SUBSTR [--kernel ...] [--workload TEXT]
      JSON of HBM bytes per launch: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of each kernel,
      scaled by what the same counters read for cpm::copy_probe_kernel, which reads and writes
      B bytes per dispatch (bytes per launch = KiB x B / copy-probe KiB)

Counter sums are per dispatch: rocprofv3 writes one row per dispatch and counter (the
per-instance values already reduced), so a kernel's value is the mean over its dispatches.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import sys
from typing import Dict, Iterable, List, Optional


def kernel_stats(path: str, top: int = 10) -> str:
    rows = list(csv.DictReader(open(path)))
    out = ["| kernel | calls | avg us | min us | max us | % time |", "|---|---:|---:|---:|---:|---:|"]
    for r in rows[:top]:
        out.append(f"| `{r['Name']}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                   f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
                   f"{float(r['Percentage']):.2f} |")
    return "\n".join(out)


def per_dispatch(path: str, match: Optional[str] = None) -> Dict[str, Dict[str, float]]:
    """{kernel: {counter: mean over dispatches of the per-dispatch value, "_dispatches": n}}."""
    sums: Dict[str, Dict[str, float]] = collections.defaultdict(lambda: collections.defaultdict(float))
    dispatches: Dict[str, set] = collections.defaultdict(set)
    for r in csv.DictReader(open(path)):
        k = r["Kernel_Name"]
        if match and match not in k:
            continue
        sums[k][r["Counter_Name"]] += float(r["Counter_Value"])
        dispatches[k].add(r["Dispatch_Id"])
    out = {}
    for k, c in sums.items():
        n = len(dispatches[k])
        out[k] = {name: v / n for name, v in c.items()}
        out[k]["_dispatches"] = n
    return out


COUNTER_COLUMNS = [
    ("SQ_INSTS_VALU", "VALU"), ("SQ_INSTS_SALU", "SALU"), ("SQ_INSTS_SMEM", "SMEM"),
    ("SQ_INSTS_VMEM_RD", "VMEM rd"), ("SQ_INSTS_VMEM_WR", "VMEM wr"), ("SQ_INSTS_LDS", "LDS"),
    ("SQ_INSTS_BRANCH", "branch"), ("SQ_WAVE_CYCLES", "wave cycles"), ("SQ_WAIT_ANY", "wait (any)"),
    ("SQ_WAIT_INST_ANY", "wait for issue"), ("SQ_ACTIVE_INST_VALU", "active VALU"),
    ("SQ_LDS_BANK_CONFLICT", "LDS bank-conflict cycles"),
]


def counters_per_wave(paths: Iterable[str], match: Optional[str] = None) -> str:
    merged: Dict[str, Dict[str, float]] = collections.defaultdict(dict)
    for p in paths:
        for k, c in per_dispatch(p, match).items():
            waves = c.get("SQ_WAVES")
            if not waves:
                continue
            merged[k]["waves/dispatch"] = waves
            for name, v in c.items():
                if name not in ("SQ_WAVES", "_dispatches"):
                    merged[k][name] = v / waves
    cols = [(key, label) for key, label in COUNTER_COLUMNS if any(key in c for c in merged.values())]
    head = "| kernel | waves/dispatch | " + " | ".join(label for _, label in cols) + " |"
    out = [head, "|---|" + "---:|" * (len(cols) + 1)]
    for k, c in merged.items():
        cells = []
        for key, _ in cols:
            v = c.get(key)
            if v is None:
                cells.append("")
            elif key in ("SQ_WAIT_ANY", "SQ_WAIT_INST_ANY") and c.get("SQ_WAVE_CYCLES"):
                cells.append(f"{v:.0f} ({100 * v / c['SQ_WAVE_CYCLES']:.0f} %)")
            else:
                cells.append(f"{v:.0f}")
        out.append(f"| `{k}` | {c['waves/dispatch']:.0f} | " + " | ".join(cells) + " |")
    return "\n".join(out)


def ratios(path: str, match: Optional[str] = None) -> str:
    data = per_dispatch(path, match)
    names = sorted({n for c in data.values() for n in c if n != "_dispatches"})
    out = ["| kernel | dispatches | " + " | ".join(names) + " |", "|---|" + "---:|" * (len(names) + 1)]
    for k, c in data.items():
        out.append(f"| `{k}` | {c['_dispatches']} | " + " | ".join(f"{c.get(n, 0):.0f}" for n in names) + " |")
    return "\n".join(out)


def traffic(fetch: str, write: str, copy_bytes: int, kernels: Dict[str, str], workload: str = "") -> dict:
    f, w = per_dispatch(fetch), per_dispatch(write)

    def find(data, sub):
        hits = [k for k in data if sub in k]
        if not hits:
            raise KeyError(f"no kernel matching {sub!r}")
        return hits[0]

    probe_f = f[find(f, "copy_probe_kernel")]["FETCH_SIZE"]
    probe_w = w[find(w, "copy_probe_kernel")]["WRITE_SIZE"]
    out = {"workload": workload,
           "calibration": {"kernel": "cpm::copy_probe_kernel", "bytes_read": copy_bytes,
                           "bytes_written": copy_bytes, "fetch_kib_per_dispatch": probe_f,
                           "write_kib_per_dispatch": probe_w,
                           "fetch_factor": probe_f * 1024 / copy_bytes,
                           "write_factor": probe_w * 1024 / copy_bytes}}
    for name, sub in kernels.items():
        kf, kw = find(f, sub), find(w, sub)
        rb = f[kf]["FETCH_SIZE"] * copy_bytes / probe_f
        wb = w[kw]["WRITE_SIZE"] * copy_bytes / probe_w
        out[name] = {"kernel": kf, "dispatches": int(f[kf]["_dispatches"]),
                     "fetch_kib": f[kf]["FETCH_SIZE"], "write_kib": w[kw]["WRITE_SIZE"],
                     "read_bytes": rb, "write_bytes": wb, "bytes_per_launch": rb + wb}
    return out


def main(argv: Optional[List[str]] = None):
    ap = argparse.ArgumentParser(prog="python -m corepotts.rocprof_summary")
    sub = ap.add_subparsers(dest="mode", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=10)
    c = sub.add_parser("counters")
    c.add_argument("csv", nargs="+")
    c.add_argument("--match")
    r = sub.add_parser("ratios")
    r.add_argument("csv")
    r.add_argument("--match")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--kernel", action="append", default=[], help="NAME=SUBSTRING")
    t.add_argument("--workload", default="")
    a = ap.parse_args(argv)
    if a.mode == "stats":
        print(kernel_stats(a.csv, a.top))
    elif a.mode == "counters":
        print(counters_per_wave(a.csv, a.match))
    elif a.mode == "ratios":
        print(ratios(a.csv, a.match))
    else:
        kernels = dict(k.split("=", 1) for k in a.kernel)
        print(json.dumps(traffic(a.fetch, a.write, a.copy_bytes, kernels, a.workload), indent=1))


if __name__ == "__main__":
    main(sys.argv[1:])
