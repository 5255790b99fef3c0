"""Summaries of rocprofv3 databases (run_results.db, rocpd SQLite schema).

    python -m admatch_amd.profsum trace <run_results.db> [--title T] [--cmd CMD]
        markdown table: calls, total and average duration, share per kernel
    python -m admatch_amd.profsum counters <run_results.db> [--kernel K] [--per N] [--per-dispatch]
        markdown table of the PMC counters summed over the matching kernels
        (and divided by N, e.g. the number of scan steps, when given), or
        one row per dispatch
    python -m admatch_amd.profsum traffic --fetch DB --write DB --copy-fetch DB
        --copy-write DB --copy-bytes B --kernel K [--skip N] [--cmd ...] > out.json
        HBM bytes per launch of kernel K from FETCH_SIZE and WRITE_SIZE (KiB
        per dispatch), each scaled by the factor that makes the same counter
        read B bytes for a plain copy of B bytes (copy_kernel)
"""
from __future__ import annotations

import argparse
import json
import sqlite3


def kernel_rows(db_path: str):
    """(name, calls, total us, average us, percent) per kernel."""
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    db.close()
    return rows


def counter_sums(db_path: str, kernel: str = ""):
    """{counter: (sum over dispatches, dispatches)} for kernels whose name contains `kernel`."""
    db = sqlite3.connect(db_path)
    rows = db.execute("select counter_name, sum(value), count(distinct dispatch_id) from counters_collection "
                      "where kernel_name like ? group by counter_name", (f"%{kernel}%",)).fetchall()
    db.close()
    return {n: (float(v), int(c)) for n, v, c in rows}


def trace(a):
    print(f"# {a.title}\n")
    if a.cmd:
        print(f"Command: `{a.cmd}`\n")
    print("| kernel | calls | total (ms) | average (ms) | share |")
    print("|---|---:|---:|---:|---:|")
    for name, calls, tot, avg, pct in kernel_rows(a.db):
        print(f"| `{name}` | {calls} | {tot / 1e3:.3f} | {avg / 1e3:.3f} | {pct:.1f}% |")


def counter_per_dispatch(db_path: str, kernel: str = ""):
    """[(dispatch id, {counter: value})] in dispatch order, for kernels whose
    name contains `kernel`."""
    db = sqlite3.connect(db_path)
    rows = db.execute("select dispatch_id, counter_name, sum(value) from counters_collection "
                      "where kernel_name like ? group by dispatch_id, counter_name order by dispatch_id",
                      (f"%{kernel}%",)).fetchall()
    db.close()
    out = {}
    for d, n, v in rows:
        out.setdefault(int(d), {})[n] = float(v)
    return sorted(out.items())


def counters(a):
    if a.per_dispatch:
        print(f"# {a.title}\n")
        if a.cmd:
            print(f"Command: `{a.cmd}`\n")
        rows = counter_per_dispatch(a.db, a.kernel)
        names = sorted({n for _, d in rows for n in d})
        print("| dispatch | " + " | ".join(names) + " |")
        print("|---:|" + "---:|" * len(names))
        for k, (d, vals) in enumerate(rows):
            print(f"| {k + 1} | " + " | ".join(f"{vals.get(n, 0):.6g}" for n in names) + " |")
        return
    print(f"# {a.title}\n")
    if a.cmd:
        print(f"Command: `{a.cmd}`\n")
    per = f" per {a.per_name}" if a.per else ""
    print(f"| counter | sum | dispatches | value{per} |")
    print("|---|---:|---:|---:|")
    for name, (v, c) in sorted(counter_sums(a.db, a.kernel).items()):
        q = v / a.per if a.per else v / max(c, 1)
        print(f"| {name} | {v:.6g} | {c} | {q:.4g} |")


def traffic(a):
    def per_dispatch(db, counter, kernel, skip=0):
        rows = [d[counter] for _, d in counter_per_dispatch(db, kernel) if counter in d][skip:]
        return sum(rows) / len(rows)
    cf = per_dispatch(a.copy_fetch, "FETCH_SIZE", "copy_kernel")
    cw = per_dispatch(a.copy_write, "WRITE_SIZE", "copy_kernel")
    f_scale = a.copy_bytes / (cf * 1024.0)
    w_scale = a.copy_bytes / (cw * 1024.0)
    ef = per_dispatch(a.fetch, "FETCH_SIZE", a.kernel, a.skip)
    ew = per_dispatch(a.write, "WRITE_SIZE", a.kernel, a.skip)
    out = {
        "kernel": a.kernel,
        "dispatches_skipped": a.skip,
        "fetch_kib_per_dispatch": ef,
        "write_kib_per_dispatch": ew,
        "copy_bytes": a.copy_bytes,
        "copy_fetch_kib": cf,
        "copy_write_kib": cw,
        "fetch_scale": f_scale,
        "write_scale": w_scale,
        "hbm_read_bytes_per_launch": ef * 1024.0 * f_scale,
        "hbm_write_bytes_per_launch": ew * 1024.0 * w_scale,
        "hbm_bytes_per_launch": ef * 1024.0 * f_scale + ew * 1024.0 * w_scale,
        "commands": a.cmd or [],
    }
    print(json.dumps(out, indent=1))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--title", default="rocprofv3 kernel summary")
    t.add_argument("--cmd", default="")
    c = sub.add_parser("counters")
    c.add_argument("db")
    c.add_argument("--kernel", default="")
    c.add_argument("--per", type=float, default=0.0)
    c.add_argument("--per-name", default="step")
    c.add_argument("--title", default="rocprofv3 counters")
    c.add_argument("--cmd", default="")
    c.add_argument("--per-dispatch", action="store_true", help="one row per dispatch instead of sums")
    r = sub.add_parser("traffic")
    r.add_argument("--fetch", required=True)
    r.add_argument("--write", required=True)
    r.add_argument("--copy-fetch", required=True)
    r.add_argument("--copy-write", required=True)
    r.add_argument("--copy-bytes", type=float, required=True)
    r.add_argument("--kernel", required=True)
    r.add_argument("--skip", type=int, default=0,
                   help="leave out the first dispatches of the kernel (the first launch after the table uploads "
                        "also writes back what the uploads left in the caches)")
    r.add_argument("--cmd", action="append")
    a = ap.parse_args(argv)
    {"trace": trace, "counters": counters, "traffic": traffic}[a.what](a)


if __name__ == "__main__":
    main()
