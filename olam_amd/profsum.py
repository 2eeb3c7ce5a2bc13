"""Summaries of rocprofv3 SQLite output (kernel times and HBM counters).

    python -m olam_amd.profsum trace  <trace_results.db>
    python -m olam_amd.profsum counter <pmc_results.db> [--per <name=cell_levels>...]

``trace`` prints per-kernel call counts, total and average duration (the
``top_kernels`` view).  ``counter`` prints the average value per dispatch of each
collected counter (FETCH_SIZE / WRITE_SIZE are in KiB) per kernel.
"""

from __future__ import annotations

import re
import sqlite3
import sys
from collections import defaultdict


def short(name):
    m = re.search(r"(k_[a-z_0-9]+)", name)
    return m.group(1) if m else name.split("(")[0][:40]


def trace_summary(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, total_calls, total_duration, average from top_kernels").fetchall()
    agg = defaultdict(lambda: [0, 0.0])
    for name, calls, tot, avg in rows:
        k = short(name)
        agg[k][0] += calls
        agg[k][1] += tot
    out = []
    for k, (calls, tot) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append((k, calls, tot / 1e6, tot / calls / 1e6))
    return out


def counter_summary(db):
    con = sqlite3.connect(db)
    rows = con.execute(
        "select k.name, p.counter_name, p.counter_value from pmc_events p "
        "join kernels k on k.dispatch_id = p.dispatch_id").fetchall()
    agg = defaultdict(lambda: [0, 0.0])
    for name, cname, val in rows:
        key = (short(name), cname)
        agg[key][0] += 1
        agg[key][1] += float(val)
    return {k: (n, v / n) for k, (n, v) in agg.items()}


def main(argv):
    mode, db = argv[1], argv[2]
    if mode == "trace":
        print(f"{'kernel':20s} {'calls':>6s} {'total_ms':>10s} {'avg_ms':>9s}")
        for k, calls, tot, avg in trace_summary(db):
            print(f"{k:20s} {calls:6d} {tot:10.3f} {avg:9.4f}")
    else:
        print(f"{'kernel':20s} {'counter':12s} {'dispatches':>10s} {'avg_per_dispatch':>18s}")
        for (k, c), (n, v) in sorted(counter_summary(db).items()):
            print(f"{k:20s} {c:12s} {n:10d} {v:18.1f}")


if __name__ == "__main__":
    main(sys.argv)
