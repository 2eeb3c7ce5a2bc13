"""Summaries of rocprofv3 SQLite output (kernel times and HBM counters).

    python -m olam_amd.profsum trace  <trace_results.db>
    python -m olam_amd.profsum counter <pmc_results.db>
    python -m olam_amd.profsum hbm <fetch.db> <write.db> <copy_fetch.db> <copy_write.db> <copy_bytes> [trace.db]

``trace`` prints per-kernel call counts, total and average duration (the
``top_kernels`` view).  ``counter`` prints the average value per dispatch of each
collected counter (FETCH_SIZE / WRITE_SIZE are in KiB) per kernel.  ``hbm`` prints a
markdown table of HBM bytes read and written per dispatch, each counter scaled by what
it reads for a plain copy of copy_bytes (read once, written once), with the average
duration and bandwidth per kernel when a kernel trace of the same command is given.
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
    """(kernel, calls, total ms, average ms) from the dispatch records (ns timestamps)."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), sum(end - start) from kernels group by name").fetchall()
    agg = defaultdict(lambda: [0, 0.0])
    for name, calls, tot in rows:
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


def hbm_table(fetch_db, write_db, cal_fetch_db, cal_write_db, copy_bytes, trace_db=None):
    f = counter_summary(fetch_db)
    w = counter_summary(write_db)
    cf = [v for (k, c), (n, v) in counter_summary(cal_fetch_db).items() if c.startswith("FETCH_SIZE")]
    cw = [v for (k, c), (n, v) in counter_summary(cal_write_db).items() if c.startswith("WRITE_SIZE")]
    kib = copy_bytes / 1024.0
    sf = kib / max(cf) if cf else 1.0
    sw = kib / max(cw) if cw else 1.0
    avg = {k: a for k, _, _, a in trace_summary(trace_db)} if trace_db else {}
    rows = []
    for (k, c), (n, v) in f.items():
        if not c.startswith("FETCH_SIZE"):
            continue
        wv = next((vv for (kk, cc), (_, vv) in w.items() if kk == k and cc.startswith("WRITE_SIZE")), 0.0)
        rg, wg = v * sf * 1024 / 1e9, wv * sw * 1024 / 1e9
        ms = avg.get(k)
        rows.append((k, v, rg, wv, wg, ms, (rg + wg) / (ms * 1e-3) if ms else None))
    rows.sort(key=lambda r: -(r[2] + r[4]))
    out = [f"Calibration: a plain copy of {copy_bytes} bytes reads FETCH_SIZE x {sf:.3f} and WRITE_SIZE x "
           f"{sw:.3f} of its bytes (factors applied below).", "",
           "| kernel | FETCH_SIZE KiB | read GB | WRITE_SIZE KiB | written GB | avg ms | GB/s |",
           "|---|---|---|---|---|---|---|"]
    for k, fv, rg, wv, wg, ms, bw in rows:
        out.append(f"| {k} | {fv:,.0f} | {rg:.2f} | {wv:,.0f} | {wg:.2f} | "
                   f"{'' if ms is None else f'{ms:.3f}'} | {'' if bw is None else f'{bw:.0f}'} |")
    return "\n".join(out)


def main(argv):
    mode, db = argv[1], argv[2]
    if mode == "hbm":
        print(hbm_table(argv[2], argv[3], argv[4], argv[5], int(argv[6]), argv[7] if len(argv) > 7 else None))
        return
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
