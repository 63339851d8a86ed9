"""Summarise rocprofv3 output into the short text kept under profiles/.

    python -m gftevolver_amd.profsummary stats DIR            # --kernel-trace --stats
    python -m gftevolver_amd.profsummary pmc DIR COUNTER      # --pmc COUNTER
    python -m gftevolver_amd.profsummary tail DIR NAME K      # --kernel-trace

`stats` prints calls / average / total per kernel from *_kernel_stats.csv;
`pmc` averages a counter per dispatch of each kernel from
*_counter_collection.csv (KiB per dispatch for FETCH_SIZE and WRITE_SIZE).
`tail` averages the duration of the last K dispatches of every kernel whose
name contains NAME, from *_kernel_trace.csv (the timed steps of a run whose
earlier dispatches of the same kernel were tuning or warm-up).
"""
from __future__ import annotations

import csv
import glob
import os
import sys
from collections import defaultdict


def _find(d, pattern):
    hits = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not hits:
        raise SystemExit(f"no {pattern} under {d}")
    return hits


def stats(d):
    rows = []
    for path in _find(d, "*kernel_stats.csv"):
        with open(path) as f:
            rows += list(csv.DictReader(f))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print(f"{'calls':>6} {'avg_us':>10} {'total_ms':>10} {'pct':>6}  kernel")
    for r in rows:
        print(f"{int(r['Calls']):>6} {float(r['AverageNs']) / 1e3:>10.2f} {float(r['TotalDurationNs']) / 1e6:>10.3f} "
              f"{float(r['Percentage']):>6.2f}  {r['Name'][:140]}")


def pmc(d, counter):
    vals = defaultdict(list)
    for path in _find(d, "*counter_collection.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") == counter:
                    vals[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    print(f"{counter}: {'dispatches':>10} {'avg_per_dispatch':>18}  kernel")
    for k, v in sorted(vals.items(), key=lambda kv: -sum(kv[1])):
        print(f"{counter}: {len(v):>10} {sum(v) / len(v):>18.1f}  {k[:140]}")


def tail(d, name, k):
    runs = defaultdict(list)
    for path in _find(d, "*kernel_trace.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                if name in r["Kernel_Name"]:
                    runs[r["Kernel_Name"]].append(r)
    if not runs:
        raise SystemExit(f"no dispatch of a kernel named like {name!r} under {d}")
    print(f"{'of_calls':>8} {'last':>5} {'avg_us':>10} {'min_us':>10} {'max_us':>10}  kernel")
    for kname, rows in runs.items():
        rows.sort(key=lambda r: int(r["Start_Timestamp"]))
        last = rows[-k:]
        us = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in last]
        print(f"{len(rows):>8} {len(last):>5} {sum(us) / len(us):>10.2f} {min(us):>10.2f} {max(us):>10.2f} "
              f" {kname[:140]}")


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "stats":
        stats(sys.argv[2])
    elif len(sys.argv) >= 4 and sys.argv[1] == "pmc":
        pmc(sys.argv[2], sys.argv[3])
    elif len(sys.argv) >= 5 and sys.argv[1] == "tail":
        tail(sys.argv[2], sys.argv[3], int(sys.argv[4]))
    else:
        raise SystemExit(__doc__)
