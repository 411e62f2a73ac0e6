"""Summarise a rocprofv3 results database (``*_results.db``) as text: per
kernel call count, total and average duration, share of GPU time.

    python -m tmp_amd.profsummary <results.db> [--title T] [--command C]
"""
from __future__ import annotations

import argparse
import sqlite3


def summary(db: str) -> list[tuple]:
    c = sqlite3.connect(db)
    rows = list(c.execute("select name, total_calls, total_duration, average, percentage from top_kernels"))
    c.close()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--title", default="rocprofv3 kernel summary")
    ap.add_argument("--command", default="")
    a = ap.parse_args()
    print(f"# {a.title}\n")
    if a.command:
        print(f"Command: `{a.command}`\n")
    print("| kernel | calls | total us | avg us | % GPU time |")
    print("|---|---:|---:|---:|---:|")
    for name, n, tot, avg, pct in summary(a.db):
        print(f"| `{name}` | {n} | {tot:.1f} | {avg:.3f} | {pct:.2f} |")


if __name__ == "__main__":
    main()
