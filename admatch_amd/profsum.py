"""Summarise a rocprofv3 kernel-trace database as markdown.

    python -m admatch_amd.profsum <run_results.db> [--title T] [--cmd CMD]

Prints one table row per kernel: calls, total and average duration, share.
"""
from __future__ import annotations

import argparse
import sqlite3


def kernel_rows(db_path: str):
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    db.close()
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--title", default="rocprofv3 kernel summary")
    ap.add_argument("--cmd", default="")
    a = ap.parse_args(argv)
    print(f"# {a.title}\n")
    if a.cmd:
        print(f"Command: `{a.cmd}`\n")
    print("| kernel | calls | total (us) | average (us) | share |")
    print("|---|---:|---:|---:|---:|")
    for name, calls, tot, avg, pct in kernel_rows(a.db):
        print(f"| `{name}` | {calls} | {tot / 1e3:.1f} | {avg / 1e3:.1f} | {pct:.1f}% |")


if __name__ == "__main__":
    main()
