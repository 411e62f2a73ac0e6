"""Summarise a rocprofv3 results database (``*_results.db``) as text: per
kernel call count, total and average duration, share of GPU time.

    python -m tmp_amd.profsummary <results.db> [--title T] [--command C]
    python -m tmp_amd.profsummary --pmc <results.db>
    python -m tmp_amd.profsummary --traffic KERNEL --cells NC --copy-bytes B \
        <fetch.db> <write.db> <copy_fetch.db> <copy_write.db>

The traffic mode turns FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of one
kernel into HBM bytes per launch, each counter scaled by the factor that makes
it read B bytes for the plain copy kernel (which moves B bytes each way).
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


def _one(db: str, kernel: str, counter: str) -> float:
    cs = counters(db)
    hits = [v[counter] for k, v in cs.items() if kernel in k and counter in v]
    if len(hits) != 1:
        raise SystemExit(f"{db}: {len(hits)} kernels match {kernel!r} with {counter}")
    return hits[0]


def traffic(kernel: str, cells: int, copy_bytes: int, dbs: list, command: str) -> None:
    fetch_db, write_db, cfetch_db, cwrite_db = dbs
    f_kib = _one(fetch_db, kernel, "FETCH_SIZE")
    w_kib = _one(write_db, kernel, "WRITE_SIZE")
    cf_kib = _one(cfetch_db, "k_copy", "FETCH_SIZE")
    cw_kib = _one(cwrite_db, "k_copy", "WRITE_SIZE")
    sf = copy_bytes / (cf_kib * 1024.0)
    sw = copy_bytes / (cw_kib * 1024.0)
    tb = f_kib * 1024.0 * sf + w_kib * 1024.0 * sw
    print(f"# HBM traffic of {kernel} per launch\n")
    if command:
        print(f"Commands: {command}\n")
    print("| quantity | value |")
    print("|---|---:|")
    print(f"| {kernel} FETCH_SIZE (KiB/dispatch) | {f_kib:.1f} |")
    print(f"| {kernel} WRITE_SIZE (KiB/dispatch) | {w_kib:.1f} |")
    print(f"| k_copy FETCH_SIZE (KiB/dispatch, {copy_bytes} B copied) | {cf_kib:.1f} |")
    print(f"| k_copy WRITE_SIZE (KiB/dispatch) | {cw_kib:.1f} |")
    print(f"| FETCH scale factor | {sf:.4f} |")
    print(f"| WRITE scale factor | {sw:.4f} |")
    print(f"| read bytes per launch | {f_kib * 1024.0 * sf:.4g} |")
    print(f"| written bytes per launch | {w_kib * 1024.0 * sw:.4g} |")
    print()
    print(f"cells: {cells}")
    print(f"traffic_bytes_per_launch: {tb:.6g}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db", nargs="+")
    ap.add_argument("--traffic", metavar="KERNEL")
    ap.add_argument("--cells", type=int, default=0)
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--title", default="rocprofv3 kernel summary")
    ap.add_argument("--command", default="")
    ap.add_argument("--pmc", action="store_true", help="the db holds --pmc counters")
    a = ap.parse_args()
    if a.traffic:
        traffic(a.traffic, a.cells, a.copy_bytes, a.db, a.command)
        return
    a.db = a.db[0]
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
