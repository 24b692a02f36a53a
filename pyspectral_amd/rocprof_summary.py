"""Summaries of rocprofv3 CSV output as Markdown tables (used for the files under profiles/).

    python -m pyspectral_amd.rocprof_summary trace RUN_kernel_trace.csv [--active-us 20]
    python -m pyspectral_amd.rocprof_summary traffic FETCH_counter_collection.csv \
        WRITE_counter_collection.csv [--kernels SUBSTR,...] [--active-frac F]
    python -m pyspectral_amd.rocprof_summary counters RUN_counter_collection.csv --ref COUNTER \
        [--kernels SUBSTR,...]

``trace``: per kernel name, the number of dispatches, how many ran longer than --active-us
(the device loop's chunks past the end of a run are no-op dispatches of a few us), their mean,
min and max duration and the total time.

``counters``: per kernel, the mean of every counter of a --pmc run over the dispatches the
``--ref`` counter marks active.

``traffic``: per kernel, the mean FETCH_SIZE and WRITE_SIZE counter values (KiB per dispatch)
over its dispatches that moved data, scaled into bytes by a calibration from the same runs: a
plain device copy of known size (``copyBuffer``; bench.py copies 1 GiB into another 1 GiB 13
times when it measures the copy bandwidth) reads and writes exactly its size, so reads =
FETCH_SIZE x (2^30 / its FETCH_SIZE) and writes likewise.  With --active-frac F only the
dispatches reaching F times the kernel's largest value count (an autotuner's timings of smaller
launches of the same kernel are left out).
"""

from __future__ import annotations

import argparse
import csv
import json
import sys
from collections import defaultdict
from pathlib import Path


def _short(name: str, width: int = 60) -> str:
    name = name.replace("void ", "")
    cut = name.find("(")
    if cut > 0:
        name = name[:cut]
    return name[:width]


def kernel_times(path: Path, active_us: float) -> list[dict]:
    rows: dict[str, list[float]] = defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3  # ns -> us
            rows[_short(r["Kernel_Name"])].append(d)
    total = sum(sum(v) for v in rows.values())
    out = []
    for name, ds in rows.items():
        act = [d for d in ds if d > active_us]
        out.append({"kernel": name, "calls": len(ds), "active": len(act),
                    "avg_us": sum(act) / len(act) if act else sum(ds) / len(ds),
                    "min_us": min(act or ds), "max_us": max(act or ds),
                    "total_ms": sum(ds) / 1e3, "pct": 100 * sum(ds) / total})
    return sorted(out, key=lambda r: -r["total_ms"])


def trace_table(path: Path, active_us: float) -> str:
    lines = [f"| kernel | calls | active (> {active_us:g} us) | active avg us | min us | "
             "max us | total ms | % |", "|---|---|---|---|---|---|---|---|"]
    for r in kernel_times(path, active_us):
        lines.append(f"| `{r['kernel']}` | {r['calls']} | {r['active']} | {r['avg_us']:.1f} | "
                     f"{r['min_us']:.1f} | {r['max_us']:.1f} | {r['total_ms']:.2f} | "
                     f"{r['pct']:.2f} |")
    return "\n".join(lines)


def counter_means(path: Path, counter: str, floor_kib: float = 64.0) -> dict[str, list[float]]:
    vals: dict[str, list[float]] = defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if r.get("Counter_Name", counter) != counter:
                continue
            v = float(r["Counter_Value"])
            if v > floor_kib:
                vals[_short(r["Kernel_Name"])].append(v)
    return vals


def _calibration(vals: dict[str, list[float]]) -> float:
    """Counter value of the copy-bandwidth copies: bench.py copies the same 1 GiB buffer 13
    times (3 untimed, 10 timed), so among the large copyBuffer dispatches it is the group of
    13 equal values (to 1 %); other large copies of a run (state loads) come in other counts."""
    big = [v for k, vs in vals.items() if "copyBuffer" in k for v in vs if v > 2.0 ** 16]
    groups: dict[float, list[float]] = defaultdict(list)
    for v in big:
        groups[float(f"{v:.2g}")].append(v)
    cands = [g for g in groups.values() if len(g) == 13]
    if not cands:
        raise SystemExit("no group of 13 equal copies (bench.py's copy bandwidth) to "
                         "calibrate on")
    best = max(cands, key=lambda g: g[0])
    return sorted(best)[len(best) // 2]


def _active(vals: dict[str, list[float]], frac: float) -> dict[str, list[float]]:
    """Per kernel, the dispatches whose value reaches ``frac`` of the kernel's largest (with
    frac > 0: drops the smaller launches of the same kernel, such as an autotuner's timings of
    sub-ensembles, from the mean)."""
    if frac <= 0:
        return vals
    return {k: [v for v in vs if v >= frac * max(vs)] for k, vs in vals.items()}


def traffic(fetch: Path, write: Path, kernels: list[str] | None,
            active_frac: float = 0.0) -> dict:
    fv, wv = counter_means(fetch, "FETCH_SIZE"), counter_means(write, "WRITE_SIZE")
    cal_f, cal_w = _calibration(fv), _calibration(wv)
    fv, wv = _active(fv, active_frac), _active(wv, active_frac)
    ff = 2.0 ** 30 / (cal_f * 1024.0)
    fw = 2.0 ** 30 / (cal_w * 1024.0)
    res = {"fetch_factor": ff, "write_factor": fw, "kernels": {}}
    for k in sorted(set(fv) | set(wv)):
        if kernels and not any(s in k for s in kernels):
            continue
        rd = sum(fv.get(k, [0.0])) / max(1, len(fv.get(k, []))) * 1024.0 * ff
        wr = sum(wv.get(k, [0.0])) / max(1, len(wv.get(k, []))) * 1024.0 * fw
        res["kernels"][k] = {"dispatches": len(fv.get(k, [])), "read_bytes": rd,
                             "write_bytes": wr, "bytes": rd + wr}
    return res


def counters(path: Path, kernels: list[str], ref: str) -> dict[str, dict[str, float]]:
    """Per kernel, the mean of every counter of a --pmc run over the kernel's active dispatches:
    those whose ``ref`` counter reaches half its largest value for that kernel (the device
    loop's no-op dispatches past the end of a run do next to nothing)."""
    by: dict[str, dict[str, dict[str, float]]] = defaultdict(lambda: defaultdict(dict))
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            k = _short(r["Kernel_Name"], 80)
            if kernels and not any(x in k for x in kernels):
                continue
            by[k][r["Dispatch_Id"]][r["Counter_Name"]] = float(r["Counter_Value"])
    out = {}
    for k, disp in by.items():
        top = max(d.get(ref, 0.0) for d in disp.values())
        act = [d for d in disp.values() if d.get(ref, 0.0) >= 0.5 * top]
        names = sorted({c for d in act for c in d})
        out[k] = {"dispatches": len(act),
                  **{c: sum(d.get(c, 0.0) for d in act) / len(act) for c in names}}
    return out


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("csv", type=Path)
    t.add_argument("--active-us", type=float, default=20.0)
    p = sub.add_parser("traffic")
    p.add_argument("fetch", type=Path)
    p.add_argument("write", type=Path)
    p.add_argument("--kernels", default="")
    p.add_argument("--active-frac", type=float, default=0.0,
                   help="keep a kernel's dispatches reaching this fraction of its largest")
    q = sub.add_parser("counters")
    q.add_argument("csv", type=Path)
    q.add_argument("--ref", required=True, help="counter that marks a dispatch active")
    q.add_argument("--kernels", default="")
    a = ap.parse_args(argv)
    if a.cmd == "counters":
        print(json.dumps(counters(a.csv, [s for s in a.kernels.split(",") if s], a.ref),
                         indent=1))
        return 0
    if a.cmd == "trace":
        print(trace_table(a.csv, a.active_us))
        return 0
    res = traffic(a.fetch, a.write, [s for s in a.kernels.split(",") if s], a.active_frac)
    print(f"calibration: reads = FETCH_SIZE x {res['fetch_factor']:.4f}, writes = WRITE_SIZE x "
          f"{res['write_factor']:.4f}\n")
    print("| kernel | dispatches | HBM read (MB) | HBM write (MB) | total (MB) |")
    print("|---|---|---|---|---|")
    for k, v in res["kernels"].items():
        print(f"| `{k}` | {v['dispatches']} | {v['read_bytes'] / 1e6:.1f} | "
              f"{v['write_bytes'] / 1e6:.1f} | {v['bytes'] / 1e6:.1f} |")
    print("\n" + json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
