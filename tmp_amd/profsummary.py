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


def counters(db: str) -> dict:
    """{kernel name: {counter: mean value per dispatch}} from a --pmc run."""
    c = sqlite3.connect(db)
    rows = list(c.execute("select kernel_name, counter_name, dispatch_id, value from counters_collection"))
    c.close()
    acc: dict = {}
    for name, cn, d, v in rows:
        acc.setdefault(name, {}).setdefault(cn, {})
        acc[name][cn][d] = acc[name][cn].get(d, 0.0) + v
    return {k: {cn: sum(dv.values()) / len(dv) for cn, dv in cs.items()} for k, cs in acc.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--title", default="rocprofv3 kernel summary")
    ap.add_argument("--command", default="")
    ap.add_argument("--pmc", action="store_true", help="the db holds --pmc counters")
    a = ap.parse_args()
    if a.pmc:
        print(f"# {a.title}\n")
        if a.command:
            print(f"Command: `{a.command}`\n")
        cs = counters(a.db)
        names = sorted({cn for v in cs.values() for cn in v})
        print("| kernel | " + " | ".join(names) + " |")
        print("|---|" + "---:|" * len(names))
        for k, v in sorted(cs.items()):
            print(f"| `{k[:60]}` | " + " | ".join(f"{v.get(n, float('nan')):.4g}" for n in names) + " |")
        return
    print(f"# {a.title}\n")
    if a.command:
        print(f"Command: `{a.command}`\n")
    print("| kernel | calls | total us | avg us | % GPU time |")
    print("|---|---:|---:|---:|---:|")
    for name, n, tot, avg, pct in summary(a.db):
        print(f"| `{name}` | {n} | {tot:.1f} | {avg:.3f} | {pct:.2f} |")


if __name__ == "__main__":
    main()
