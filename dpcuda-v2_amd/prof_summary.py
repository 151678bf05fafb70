#!/usr/bin/env python3
"""Summarise rocprofv3 SQLite results (``*_results.db``) as short text tables.

    python3 prof_summary.py trace DIR [--top K]
    python3 prof_summary.py pmc DIR [--kernel SUBSTRING]
    python3 prof_summary.py gbs DIR --bytes SUBSTRING=BYTES [...] [--min-us T]

``trace``: per kernel name, dispatch count, mean / total / min / max duration (from
``rocprofv3 --kernel-trace --stats``).  ``pmc``: per kernel and counter, dispatch
count, mean and median of the counter value (from ``rocprofv3 --pmc ...``); with
FETCH_SIZE / WRITE_SIZE the values are KiB per dispatch.  ``gbs``: for each kernel whose
name contains SUBSTRING, the dispatches that ran longer than T microseconds (launches of a
gated kernel that only skip are shorter) and the algorithmic GB/s of BYTES per dispatch over
their mean duration.
"""
from __future__ import annotations

import argparse
import glob
import os
import sqlite3
import statistics


def _dbs(root):
    return sorted(glob.glob(os.path.join(root, "**", "*_results.db"), recursive=True))


def trace(root, top):
    rows = {}
    for db in _dbs(root):
        con = sqlite3.connect(db)
        for name, dur in con.execute("select name, duration from kernels"):
            rows.setdefault(name, []).append(dur)
        con.close()
    out = [f"{'kernel':70s} {'n':>7s} {'mean_us':>9s} {'total_ms':>9s} {'min_us':>8s} "
           f"{'max_us':>8s}"]
    order = sorted(rows.items(), key=lambda kv: -sum(kv[1]))[:top]
    for name, d in order:
        out.append(f"{name[:70]:70s} {len(d):7d} {sum(d) / len(d) / 1e3:9.3f} "
                   f"{sum(d) / 1e6:9.3f} {min(d) / 1e3:8.3f} {max(d) / 1e3:8.3f}")
    return "\n".join(out)


def pmc(root, kernel):
    vals = {}
    for db in _dbs(root):
        con = sqlite3.connect(db)
        cols = [r[1] for r in con.execute("pragma table_info(counters_collection)")]
        kn = "kernel_name" if "kernel_name" in cols else "name"
        for name, cn, v in con.execute(
                f"select {kn}, counter_name, value from counters_collection"):
            if kernel and kernel not in name:
                continue
            vals.setdefault((name, cn), []).append(v)
        con.close()
    out = [f"{'kernel':60s} {'counter':14s} {'n':>6s} {'mean':>14s} {'median':>14s}"]
    for (name, cn), v in sorted(vals.items()):
        out.append(f"{name[:60]:60s} {cn:14s} {len(v):6d} {sum(v) / len(v):14.1f} "
                   f"{statistics.median(v):14.1f}")
    return "\n".join(out)


def gbs(root, specs, min_us):
    rows = {}
    for db in _dbs(root):
        con = sqlite3.connect(db)
        for name, dur in con.execute("select name, duration from kernels"):
            rows.setdefault(name, []).append(dur)
        con.close()
    out = [f"{'kernel':52s} {'bytes':>12s} {'n_run':>6s} {'mean_us':>9s} {'GB/s':>8s} "
           f"{'of 8 TB/s':>9s}"]
    for spec in specs:
        sub, b = spec.split("=")
        nbytes = float(b)
        for name, d in sorted(rows.items()):
            if sub not in name:
                continue
            run = [x for x in d if x > min_us * 1e3]
            if not run:
                continue
            mean = sum(run) / len(run)
            g = nbytes / mean  # bytes per ns = GB/s
            out.append(f"{name[:52]:52s} {nbytes:12.4g} {len(run):6d} {mean / 1e3:9.3f} "
                       f"{g:8.1f} {g / 8000:9.3f}")
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=("trace", "pmc", "gbs"))
    ap.add_argument("dir")
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--kernel", default="")
    ap.add_argument("--bytes", nargs="*", default=[])
    ap.add_argument("--min-us", type=float, default=3.0)
    a = ap.parse_args()
    if a.mode == "gbs":
        print(gbs(a.dir, a.bytes, a.min_us))
    else:
        print(trace(a.dir, a.top) if a.mode == "trace" else pmc(a.dir, a.kernel))


if __name__ == "__main__":
    main()
